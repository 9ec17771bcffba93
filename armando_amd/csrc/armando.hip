// armando.hip -- C ABI (include/armando.h) over the gfx950 WCSPH kernels.
//
// One context owns the particle state in HBM (structure of arrays, sorted by
// (cell, id) after every link_list), the sort workspace and one HIP stream.
// The public entry points mirror the reference's routines; the fused
// arm_time_integration runs the same kernels back to back on the stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/armando.h"
#include "radix_sort.h"
#include "cell_sort.h"
#include "sph_kernels.h"
#include "sph_tiled.h"
#include "sph_adjoint.h"
#include "sph3d.h"
#include "sph3d_adjoint.h"
#include "sph_ghost.h"

using namespace arm;

namespace {

struct DevBuf {
  void* p = nullptr;
};

template <typename R> struct Layout {
  static constexpr bool f32 = sizeof(R) == 4;
};

}  // namespace

struct arm_ctx {
  arm_params prm{};
  KParams K{};
  int cap = 0;
  int n = 0, np = 0;
  // mirror ghosts (ARM_GHOSTS): particles nb = np + nv onward are ghosts; ng
  // of them live in the sorted state; after arm_ghost_particles the slots
  // from gfirst hold ng_new new ones and those below are stale until the
  // next arm_link_list (ghost_pending)
  bool ghosts = false;
  int nb = 0, ng = 0, ng_new = 0, gfirst = 0;
  bool ghost_pending = false;
  bool have_ranges = false;  // cstart describes the current slot order
  Ghosts G{};
  int* d_ng = nullptr;   // = d_err + 2: read back with the error counts
  int* h_pin = nullptr;  // pinned host copy of d_err[0..2]
  bool loaded = false, sorted = false, have_rates = false;
  bool tiled = true;  // LDS-tiled gathers (false: direct global-memory gathers)
  struct AdjRun* adj = nullptr;  // checkpoints of the last arm_rollout_forward
  int objkind = ARM_OBJ_PRESSURE;  // objective of the next rollout
  double grads[ARM_NGRAD] = {0};   // parameter gradients of the last backward
  bool masks_valid = false;  // nbm holds the neighbour masks of the current order
  uint4* nbm = nullptr;
  // 2D tiled gathers: the row pieces of the current order (RowPieces)
  int2* pieces = nullptr;
  int* npieces = nullptr;
  bool pieces_valid = false;
  // deposition heat source (arm_set_deposition): the map on the device, the
  // deposition time and the number of the step the next force pass belongs
  // to (its of single_step.f:219)
  double* d_dep = nullptr;
  double dep_tend = 0.0;
  int its_next = 1;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // state double buffer (the active set is A)
  St<double> A64{}, B64{};
  St<float> A32{}, B32{};
  // three dimensions (prm.dim == 3): the 3D state double buffer
  St3<double> A3_64{}, B3_64{};
  St3<float> A3_32{}, B3_32{};
  uint32_t* nbm3 = nullptr;  // 3D neighbour masks (M3_TOTAL words per particle)
  void* pack3_vq = nullptr;  // 3D force pass: packed neighbour records (Pack3)
  void* d_cd = nullptr;      // ARM_CONTACT: contact displacement per slot (V2<R>)
  // sort workspace
  uint32_t *key_by_id = nullptr, *slot_by_id = nullptr;
  uint32_t *sk[2] = {nullptr, nullptr}, *sv[2] = {nullptr, nullptr};
  uint32_t* hist = nullptr;
  uint32_t* tot = nullptr;
  int ntiles = 0, sort_bits = 0, sort_passes = 0;
  // counting sort over cells (grids of up to CS_TILE * CS_MAX_TILES cells)
  bool counting = false;
  int cs_tiles = 0;
  int* cnt = nullptr;       // per-cell counters, zero between sorts
  int* tsum = nullptr;      // per-tile sums of the counters
  uint64_t* tstat = nullptr;  // status words of the single-pass scan (k_cs_scan1)
  uint32_t scan_epoch = 0;    // epoch of the last single-pass scan
  uint64_t* ent = nullptr;  // entries grouped by cell: cap ids then cap slots (ent_of); cells in sk[1]
  uint32_t* skey = nullptr;  // sorted keys of the current order
  int* cstart = nullptr;
  // d_err[0]: particles outside the grid, [1]: h outgrew the cell, [2]: ghost
  // count, [3]: the halt word of a fused run (KParams::halt)
  int* d_err = nullptr;
  double* d_part = nullptr;
  int* d_cnt = nullptr;
  int domain_bad = 0;
  std::string err;
  // timing: (start, stop, pass) event triples recorded on the stream and
  // resolved when the times are read, so timing adds no host synchronisation
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> ev_pool;
  struct Rec {
    hipEvent_t a, b;
    int pass;
  };
  std::vector<Rec> pending;
  double pass_ms[ARM_PASS_COUNT] = {0};
  int pass_cnt[ARM_PASS_COUNT] = {0};
};

namespace {

int fail(arm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((c), ARM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

int grid_of(int n) { return (n + BLOCK - 1) / BLOCK; }

template <typename T> int dalloc(arm_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) return fail(c, ARM_ERR_HIP, "hipMalloc(%zu) failed: %s", count * sizeof(T), hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return ARM_OK;
}

template <typename R> int alloc_state(arm_ctx* c, St<R>& S, int cap) {
  int rc = ARM_OK;
#define AL(f, cnt) \
  if ((rc = dalloc(c, &S.f, (cnt))) != ARM_OK) return rc;
  AL(id, cap);
  AL(xy, cap);
  if (sizeof(R) == 4) {
    AL(cxy, cap);
    AL(off, cap);
    AL(uf, cap);
  }
  AL(v, cap);
  AL(rho, cap);
  AL(u, cap);
  AL(p, cap);
  AL(psi, cap);
  AL(drho, cap);
  AL(du, cap);
  AL(a, cap);
  AL(av, cap);
  AL(divv, cap);
  AL(rho_min, cap);
  AL(u_min, cap);
  AL(v_min, cap);
  AL(flag, cap);
  AL(tau, 4 * (size_t)cap);
  AL(hs, cap);
  AL(hn, cap);
#undef AL
  return ARM_OK;
}

bool is32(const arm_ctx* c) { return c->prm.precision == ARM_FP32; }

// ---- timing ---------------------------------------------------------------
hipEvent_t take_event(arm_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Folds finished event pairs into the per-pass totals.
void resolve_times(arm_ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (const auto& r : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->pass_ms[r.pass] += ms;
      c->pass_cnt[r.pass] += 1;
    }
    c->ev_pool.push_back(r.a);
    c->ev_pool.push_back(r.b);
  }
  c->pending.clear();
}

struct PassTimer {
  arm_ctx* c;
  int pass;
  hipEvent_t a = nullptr;
  PassTimer(arm_ctx* c_, int p) : c(c_), pass(p) {
    if (!c->timing) return;
    if (c->pending.size() > 4096) resolve_times(c);
    a = take_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
  }
  ~PassTimer() {
    if (!c->timing || !a) return;
    hipEvent_t b = take_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({a, b, pass});
  }
};

// ---- sort -------------------------------------------------------------------
// Sorts (key_by_id, slot_by_id) into (skey, sval); returns the value array.
// Sorts (key_by_id, slot_by_id) into (cell, id) order: skey and *sval_out
// hold the sorted keys and source slots.  The counting sort also writes
// cell_start (*ranges_done = true); the radix sort leaves it to the reorder.
// Where the hashing kernels put the keys for the next sort.
KeyOut key_out(arm_ctx* c) {
  if (c->counting) return KeyOut{nullptr, nullptr, c->key_by_id, c->cnt};
  return KeyOut{c->key_by_id, c->slot_by_id, nullptr, nullptr};
}

// Cells of the counting sort: the grid, plus with ghosts the extra cell
// ncell that takes the ghosts of the previous step after every real cell.
int sort_cells(const arm_ctx* c) { return c->K.ncell + (c->ghosts ? 1 : 0); }

// Clears the counters of the counting sort before a run of keyed steps.
void zero_counts(arm_ctx* c) {
  if (c->counting)
    hipLaunchKernelGGL(k_cs_zero, dim3(std::min(1024, (sort_cells(c) + CS_THREADS - 1) / CS_THREADS)),
                       dim3(CS_THREADS), 0, c->stream, c->cnt, sort_cells(c));
}

// Sorts the first n slots (c->n when n < 0).
// The entry arrays of the counting sort: ids then slots, capacity each, in
// the context's 8-byte-per-particle buffer.
Ent ent_of(arm_ctx* c) {
  uint32_t* p = reinterpret_cast<uint32_t*>(c->ent);
  return Ent{p, p + c->cap};
}

RowPieces row_pieces_of(const arm_ctx* c) {
  return c->pieces ? RowPieces{c->pieces, c->npieces, c->K.nx, c->K.ny} : RowPieces{};
}

int run_sort(arm_ctx* c, const int32_t* id_of, const uint32_t** sval_out, bool* ranges_done,
             int n_sort = -1) {
  const int n = n_sort < 0 ? c->n : n_sort;
  if (c->counting) {
    const int ncell = sort_cells(c);
    const int nb = (n + CS_THREADS - 1) / CS_THREADS;
    const Halt hl{c->K.halt, c->K.its};
    // epochs 1 .. SCAN_EPOCH_MAX; the cleared words (epoch 0) read as unpublished
    if (c->scan_epoch == SCAN_EPOCH_MAX) {
      hipLaunchKernelGGL(k_zero_words, dim3(grid_of(2 * c->cs_tiles)), dim3(BLOCK), 0, c->stream,
                         reinterpret_cast<uint32_t*>(c->tstat), (size_t)2 * c->cs_tiles);
      c->scan_epoch = 0;
    }
    ++c->scan_epoch;
    hipLaunchKernelGGL(k_cs_scan1, dim3(c->cs_tiles), dim3(CS_THREADS), 0, c->stream, c->cnt,
                       ncell, n, c->cstart, c->tstat, c->scan_epoch, hl);
    hipLaunchKernelGGL(k_cs_scatter, dim3(nb), dim3(CS_THREADS), 0, c->stream, c->key_by_id,
                       id_of, n, c->cstart, c->cnt, ent_of(c), c->sk[1], hl);
    hipLaunchKernelGGL(k_cs_rank, dim3(nb), dim3(CS_THREADS), 0, c->stream, ent_of(c), c->sk[1],
                       n, c->cstart, c->sk[0], c->sv[0], c->ghosts ? c->K.ncell : -1, hl);
    HIPCHK(c, hipGetLastError());
    c->skey = c->sk[0];
    *sval_out = c->sv[0];
    *ranges_done = true;
    c->pieces_valid = false;
    return ARM_OK;
  }
  *ranges_done = false;
  c->pieces_valid = false;
  const int ntiles = (n + SORT_TILE - 1) / SORT_TILE;
  const uint32_t* kin = c->key_by_id;
  const uint32_t* vin = c->slot_by_id;
  int dst = 0;
  for (int ps = 0; ps < c->sort_passes; ++ps) {
    const int shift = ps * c->sort_bits;
    const int bits = std::min(c->sort_bits, std::max(1, 32 - shift));
    const int bins = 1 << bits;
    hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(SORT_THREADS), 0, c->stream, kin, n,
                       shift, bins, ntiles, c->hist);
    hipLaunchKernelGGL(k_sort_scan, dim3(bins), dim3(64), 0, c->stream, c->hist, c->tot, ntiles);
    hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(SORT_THREADS), 0, c->stream, kin,
                       vin, c->sk[dst], c->sv[dst], n, shift, bits, ntiles, c->hist, c->tot);
    kin = c->sk[dst];
    vin = c->sv[dst];
    dst ^= 1;
  }
  HIPCHK(c, hipGetLastError());
  c->skey = const_cast<uint32_t*>(kin);
  *sval_out = vin;
  return ARM_OK;
}

template <typename R> St<R>& stA(arm_ctx* c);
template <> St<double>& stA<double>(arm_ctx* c) { return c->A64; }
template <> St<float>& stA<float>(arm_ctx* c) { return c->A32; }
template <typename R> St<R>& stB(arm_ctx* c);
template <> St<double>& stB<double>(arm_ctx* c) { return c->B64; }
template <> St<float>& stB<float>(arm_ctx* c) { return c->B32; }

// Swaps the active state buffer with the other one.
void swap_state(arm_ctx* c) {
  if (c->prm.dim == 3) {
    if (c->prm.precision == ARM_FP32)
      std::swap(c->A3_32, c->B3_32);
    else
      std::swap(c->A3_64, c->B3_64);
  } else if (c->prm.precision == ARM_FP32) {
    std::swap(c->A32, c->B32);
  } else {
    std::swap(c->A64, c->B64);
  }
}

// Reads and clears the error words: d_err[0] particles outside the hash
// grid, d_err[1] particles whose evolving smoothing length outgrew the cell
// (2h > cell), d_err[3] the halt word.  its_end: one past the last step of a
// fused run whose kernels halted at the first failure (-1 otherwise).  Each
// halted step still swapped the state buffers on the host without writing
// them, so an odd number of halted steps is undone here: the state is then
// the one the failing step left.
int check_domain(arm_ctx* c, int its_end = -1) {
  int bad[4] = {0, 0, 0, 0};
  HIPCHK(c, hipMemcpyAsync(bad, c->d_err, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int stop = bad[3];
  if (stop && its_end >= 0 && (its_end - stop) % 2 != 0) swap_state(c);
  if (bad[0] || bad[1] || stop) {
    c->domain_bad = bad[0] + bad[1];
    HIPCHK(c, hipMemsetAsync(c->d_err, 0, 2 * sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_err + 3, 0, sizeof(int), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    char where[96] = "";
    if (stop && its_end >= 0)
      snprintf(where, sizeof(where), " after step %d (steps %d.. not run; the state is that of step %d)",
               stop - 1, stop, stop - 1);
    if (bad[0])
      return fail(c, ARM_ERR_DOMAIN, "%d particle(s) outside the hash grid%s", bad[0], where);
    return fail(c, ARM_ERR_DOMAIN, "%d particle(s) with 2h larger than the cell %g%s", bad[1],
                c->K.cell, where);
  }
  return ARM_OK;
}

// A fused run of steps that stops at its first domain error: while it lives
// the step kernels read the halt word (cleared at its start by a fill
// kernel) and the step number c->K.its.
struct FusedRun {
  arm_ctx* c;
  FusedRun(arm_ctx* c_, int its0) : c(c_) {
    c->K.halt = c->d_err + 3;
    c->K.its = its0 - 1;  // the keys hashed before the first step
    const int nb = 1;
    hipLaunchKernelGGL(k_zero_words, dim3(nb), dim3(BLOCK), 0, c->stream,
                       reinterpret_cast<uint32_t*>(c->K.halt), (size_t)1);
  }
  ~FusedRun() {
    c->K.halt = nullptr;
    c->K.its = 0;
  }
};

// keys of the current state -> sort -> permute (all fields) + cell ranges
// With ghosts: only the set made by arm_ghost_particles since the last sort
// (slots from gfirst) enters the sorted state; every older ghost drops out.
template <typename R> int link_list_t(arm_ctx* c) {
  const int n = c->n;
  St<R>& A = stA<R>(c);
  St<R>& B = stB<R>(c);
  const int gfirst = c->ghost_pending ? c->gfirst : n;
  {
    PassTimer tm(c, ARM_PASS_KEYS);
    zero_counts(c);
    hipLaunchKernelGGL(k_keys<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                       key_out(c), c->d_err, gfirst);
  }
  const uint32_t* sval = nullptr;
  bool ranges = false;
  {
    PassTimer tm(c, ARM_PASS_SORT);
    int rc = run_sort(c, A.id, &sval, &ranges);
    if (rc) return rc;
  }
  c->ng = c->ghost_pending ? c->ng_new : 0;
  c->ghost_pending = false;
  c->n = c->K.n = c->nb + c->ng;
  c->have_ranges = true;
  {
    PassTimer tm(c, ARM_PASS_REORDER);
    // with the sort's cell ranges the kernel also adds the row pieces
    const RowPieces rp = ranges ? row_pieces_of(c) : RowPieces{};
    hipLaunchKernelGGL(k_reorder_all<R>, dim3(grid_of(c->n) + row_blocks(rp)),
                       dim3(BLOCK), 0, c->stream, A, B, c->K, c->skey, sval, c->cstart,
                       ranges ? 0 : 1, c->G, rp);
    c->pieces_valid = rp.pieces != nullptr;
  }
  HIPCHK(c, hipGetLastError());
  std::swap(A, B);
  c->sorted = true;
  c->masks_valid = false;
  return check_domain(c);
}

// Where the tiled gathers' workgroups find their segments, and how many
// workgroups they take: row pieces (CW = 0; the piece table is rebuilt after
// each sort, and the grid is the bound n / PIECE + ny on the piece count) or
// runs of CW cells of every row.
template <int CW> SegArg seg_arg(arm_ctx* c) {
  if constexpr (CW > 0) {
    return SegArg{(c->K.nx + CW - 1) / CW, nullptr, nullptr};
  } else {
    if (!c->pieces_valid) {
      const RowPieces rp = row_pieces_of(c);
      hipLaunchKernelGGL(k_row_pieces, dim3(row_blocks(rp)), dim3(BLOCK), 0, c->stream,
                         c->cstart, rp, Halt{c->K.halt, c->K.its});
      c->pieces_valid = true;
    }
    return SegArg{0, c->pieces, c->npieces};
  }
}
template <int CW> int seg_grid(const arm_ctx* c, const SegArg& sa) {
  return CW > 0 ? c->K.ny * sa.nseg : (c->n + PIECE - 1) / PIECE + c->K.ny;
}

template <typename R, bool SUM, bool EOS> void launch_density(arm_ctx* c) {
  St<R>& A = stA<R>(c);
  using T = TileCfg<R>;
  if (c->tiled && !c->K.sle) {
    const SegArg sa = seg_arg<T::CW>(c);
    const int grid = seg_grid<T::CW>(c, sa);
    if (c->K.periodic)
      hipLaunchKernelGGL((k_density_tiled<R, SUM, EOS, true, T::BLK, T::CW, T::CAP>),
                         dim3(grid), dim3(T::BLK), 0, c->stream, A, c->K, c->skey, c->cstart, sa,
                         c->nbm);
    else
      hipLaunchKernelGGL((k_density_tiled<R, SUM, EOS, false, T::BLK, T::CW, T::CAP>),
                         dim3(grid), dim3(T::BLK), 0, c->stream, A, c->K, c->skey, c->cstart, sa,
                         c->nbm);
  } else {
    hipLaunchKernelGGL((k_density<R, SUM, EOS>), dim3(grid_of(c->n)), dim3(BLOCK), 0,
                       c->stream, A, c->K, c->skey, c->cstart, c->nbm);
  }
  c->masks_valid = true;
}

// single_step.f:188-196: after the EOS the ghosts take rho, p, u of their
// parents.
template <typename R> void ghost_copy(arm_ctx* c) {
  if (c->ghosts && c->ng > 0)
    hipLaunchKernelGGL(k_ghost_copy<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                       stA<R>(c), c->K, c->G);
}

// density pass; with eos the equation of state is applied in the same kernel
template <typename R> int density_t(arm_ctx* c, int method, bool eos = false) {
  PassTimer tm(c, ARM_PASS_DENSITY);
  if (method == 1)
    eos ? launch_density<R, true, true>(c) : launch_density<R, true, false>(c);
  else
    eos ? launch_density<R, false, true>(c) : launch_density<R, false, false>(c);
  if (eos) ghost_copy<R>(c);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

template <typename R> int eos_t(arm_ctx* c) {
  St<R>& A = stA<R>(c);
  PassTimer tm(c, ARM_PASS_EOS);
  hipLaunchKernelGGL(k_eos<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, A, c->K);
  ghost_copy<R>(c);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

// The FP32 tiled force gather of the dam-break flag set (pressure, artificial
// viscosity and heat, XSPH) with pa_sph and same_e2 fixed at compile time.
template <int PC> void launch_force_hot(arm_ctx* c, const uint4* nbm) {
  using T = HotForceCfg;
  const SegArg sa = seg_arg<T::CW>(c);
  hipLaunchKernelGGL((k_force_hot<T::BLK, T::CAP, PC>), dim3(seg_grid<T::CW>(c, sa)),
                     dim3(T::BLK), 0, c->stream, stA<float>(c), c->K, c->skey, c->cstart, sa,
                     nbm);
}

template <typename R, bool P, bool AV, bool AH, bool XS, bool VI = false>
void launch_force(arm_ctx* c) {
  St<R>& A = stA<R>(c);
  using T = ForceCfg<R, VI>;
  if (c->tiled && !c->K.sle) {
    const uint4* nbm = c->masks_valid ? c->nbm : nullptr;
    if (sizeof(R) == 4 && P && AV && AH && XS && !VI && !c->K.periodic &&
        !(c->K.flags & ARM_CONDUCTION)) {
      switch ((c->K.pa_sph == 0 ? 1 : 0) | (c->K.same_e2 ? 2 : 0)) {
        case 0: launch_force_hot<0>(c, nbm); break;
        case 1: launch_force_hot<1>(c, nbm); break;
        case 2: launch_force_hot<2>(c, nbm); break;
        default: launch_force_hot<3>(c, nbm); break;
      }
    } else {
      const SegArg sa = seg_arg<T::CW>(c);
      const int grid = seg_grid<T::CW>(c, sa);
      if (c->K.periodic)
        hipLaunchKernelGGL((k_force_tiled<R, P, AV, AH, XS, VI, true, T::BLK, T::CW, T::CAP>),
                           dim3(grid), dim3(T::BLK), 0, c->stream, A, c->K, c->skey, c->cstart,
                           sa, nbm);
      else
        hipLaunchKernelGGL((k_force_tiled<R, P, AV, AH, XS, VI, false, T::BLK, T::CW, T::CAP>),
                           dim3(grid), dim3(T::BLK), 0, c->stream, A, c->K, c->skey, c->cstart,
                           sa, nbm);
    }
  } else {
    hipLaunchKernelGGL((k_force<R, P, AV, AH, XS, VI>), dim3(grid_of(c->n)), dim3(BLOCK), 0,
                       c->stream, A, c->K, c->skey, c->cstart, c->d_err);
  }
}

// Shear tensor of the physical viscosity, after the EOS (tau of every particle).
template <typename R> void launch_tensor(arm_ctx* c) {
  St<R>& A = stA<R>(c);
  using T = TileCfg<R>;
  if (c->tiled && !c->K.sle) {
    const uint4* nbm = c->masks_valid ? c->nbm : nullptr;
    const SegArg sa = seg_arg<T::CW>(c);
    const int grid = seg_grid<T::CW>(c, sa);
    if (c->K.periodic)
      hipLaunchKernelGGL((k_tensor_tiled<R, true, T::BLK, T::CW, T::CAP>), dim3(grid),
                         dim3(T::BLK), 0, c->stream, A, c->K, c->skey, c->cstart, sa, nbm);
    else
      hipLaunchKernelGGL((k_tensor_tiled<R, false, T::BLK, T::CW, T::CAP>), dim3(grid),
                         dim3(T::BLK), 0, c->stream, A, c->K, c->skey, c->cstart, sa, nbm);
  } else {
    hipLaunchKernelGGL(k_tensor<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, A, c->K,
                       c->skey, c->cstart);
  }
}

// The deposition heat source acts in the steps with its dt <= t_end
// (single_step.f:219).
void dep_window(arm_ctx* c) {
  if (c->d_dep && c->its_next * c->prm.dt <= c->dep_tend)
    c->K.flags |= KF_DEPOSITION;
  else
    c->K.flags &= ~KF_DEPOSITION;
}

// dispatch over the compile-time flag combination
template <typename R> int force_t(arm_ctx* c, bool press, int flags, bool xs) {
  const bool av = (flags & ARM_ART_VISC) != 0, ah = (flags & ARM_ART_HEAT) != 0;
  dep_window(c);
  const bool vi = press && (flags & ARM_VISC) != 0;
  if (press && (flags & ARM_CONTACT) && !c->d_cd)
    return fail(c, ARM_ERR_ARG, "ARM_CONTACT: the context was created without it");
  PassTimer tm(c, ARM_PASS_FORCE);
  if (vi) launch_tensor<R>(c);
  const int code = (vi ? 16 : 0) | (press ? 8 : 0) | (av ? 4 : 0) | (ah ? 2 : 0) | (xs ? 1 : 0);
  switch (code) {
    case 8: launch_force<R, true, false, false, false>(c); break;
    case 9: launch_force<R, true, false, false, true>(c); break;
    case 10: launch_force<R, true, false, true, false>(c); break;
    case 11: launch_force<R, true, false, true, true>(c); break;
    case 12: launch_force<R, true, true, false, false>(c); break;
    case 13: launch_force<R, true, true, false, true>(c); break;
    case 14: launch_force<R, true, true, true, false>(c); break;
    case 15: launch_force<R, true, true, true, true>(c); break;
    case 24: launch_force<R, true, false, false, false, true>(c); break;
    case 25: launch_force<R, true, false, false, true, true>(c); break;
    case 26: launch_force<R, true, false, true, false, true>(c); break;
    case 27: launch_force<R, true, false, true, true, true>(c); break;
    case 28: launch_force<R, true, true, false, false, true>(c); break;
    case 29: launch_force<R, true, true, false, true, true>(c); break;
    case 30: launch_force<R, true, true, true, false, true>(c); break;
    case 31: launch_force<R, true, true, true, true, true>(c); break;
    case 1: launch_force<R, false, false, false, true>(c); break;
    default: return fail(c, ARM_ERR_ARG, "force pass with nothing to compute");
  }
  if (press && (c->K.flags & KF_DEPOSITION))
    hipLaunchKernelGGL(k_deposition<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                       stA<R>(c), c->K);
  if (press && (flags & ARM_CONTACT)) {
    // time_integration.f90:128-142: the push to the contact distance from
    // the virtual neighbours, smoothed over the real ones, as an acceleration
    using V = typename V2<R>::type;
    V* cd = static_cast<V*>(c->d_cd);
    hipLaunchKernelGGL(k_contact_dx<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                       stA<R>(c), c->K, c->skey, c->cstart, cd);
    hipLaunchKernelGGL(k_contact_apply<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                       stA<R>(c), c->K, c->skey, c->cstart, static_cast<const V*>(cd));
  }
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

// Clears a device array inside the step with a fill kernel (not a memset
// node, so that a caller may capture the step in a graph).
int zero_bytes(arm_ctx* c, void* p, size_t bytes) {
  const size_t nw = bytes / 4;
  const int nb = (int)std::min<size_t>((nw + BLOCK - 1) / BLOCK, 4096);
  if (nb > 0)
    hipLaunchKernelGGL(k_zero_words, dim3(nb), dim3(BLOCK), 0, c->stream,
                       static_cast<uint32_t*>(p), nw);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

template <typename R> int zero_av(arm_ctx* c) {
  return zero_bytes(c, stA<R>(c).av, sizeof(typename V2<R>::type) * (size_t)c->n);
}

// The smoothing lengths written by the force pass become the current ones
// (hsml.f: h_upgrade closes single_step).
template <typename R> void take_new_h(arm_ctx* c) {
  if (c->K.sle) std::swap(stA<R>(c).hs, stA<R>(c).hn);
}

template <typename R> int single_step_t(arm_ctx* c) {
  int rc;
  if ((rc = density_t<R>(c, c->prm.density_method, true))) return rc;
  const bool xs = (c->prm.flags & ARM_AV_VEL) != 0;
  if ((rc = force_t<R>(c, true, c->prm.flags, xs))) return rc;
  take_new_h<R>(c);
  if (!xs && (rc = zero_av<R>(c))) return rc;
  c->have_rates = true;
  return ARM_OK;
}

template <typename R> int predictor_t(arm_ctx* c) {
  St<R>& A = stA<R>(c);
  PassTimer tm(c, ARM_PASS_INTEGRATE);
  hipLaunchKernelGGL(k_predictor<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, A, c->K);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

template <typename R> int corrector_t(arm_ctx* c, int its, bool keys) {
  St<R>& A = stA<R>(c);
  c->its_next = its + 1;
  PassTimer tm(c, ARM_PASS_INTEGRATE);
  hipLaunchKernelGGL(k_corrector<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, A, c->K,
                     its, keys ? 1 : 0, keys ? key_out(c) : KeyOut{}, c->d_err);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

// time_integration.f:131-235, fused: per step
//   [keys of the current positions] -> sort -> reorder + predictor
//   -> density -> eos -> forces (+ XSPH) -> corrector + keys of the next step
// Keys of the current positions, the start of every fused run of steps.
template <typename R> int begin_steps(arm_ctx* c) {
  St<R>& A = stA<R>(c);
  PassTimer tm(c, ARM_PASS_KEYS);
  zero_counts(c);
  // every ghost present is replaced by the first step's own
  hipLaunchKernelGGL(k_keys<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, A, c->K,
                     key_out(c), c->d_err, c->n);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

// ghost.f on the sorted state: select the wall of every real particle,
// number the ghosts in parent order (a scan of the flags by id) and write
// them after the n_old live slots.  keys: emit their keys for the counting
// sort (the fused step); the per-routine link_list hashes them itself.
// Reads the ghost count back (the next launches are sized by it).
template <typename R> int make_ghosts(arm_ctx* c, bool keys) {
  St<R>& A = stA<R>(c);
  const int n_old = c->n, np = c->np;
  c->ng_new = 0;
  if (np == 0) return ARM_OK;
  PassTimer tm(c, ARM_PASS_KEYS);
  hipLaunchKernelGGL(k_ghost_select<R>, dim3(grid_of(n_old)), dim3(BLOCK), 0, c->stream, A,
                     c->K, c->G, c->skey, n_old);
  const int tiles = (np + CS_TILE - 1) / CS_TILE;
  hipLaunchKernelGGL(k_cs_tile_sum, dim3(tiles), dim3(CS_THREADS), 0, c->stream, c->G.gflag, np,
                     c->tsum, Halt{});
  hipLaunchKernelGGL(k_cs_tile_scan, dim3(tiles), dim3(CS_THREADS), 0, c->stream, c->G.gflag,
                     np, 0, c->tsum, c->G.grank, Halt{});
  hipLaunchKernelGGL(k_ghost_make<R>, dim3(grid_of(n_old)), dim3(BLOCK), 0, c->stream, A, c->K,
                     c->G, n_old, c->cap, keys ? 1 : 0, keys ? key_out(c) : KeyOut{}, c->d_err,
                     c->d_ng);
  HIPCHK(c, hipGetLastError());
  // one copy: the error counts and the ghost count (d_ng = d_err + 2)
  HIPCHK(c, hipMemcpyAsync(c->h_pin, c->d_err, 3 * sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int hv[3] = {c->h_pin[2], c->h_pin[0], c->h_pin[1]};
  if ((long long)n_old + hv[0] > c->cap) {
    c->loaded = false;
    return fail(c, ARM_ERR_CAPACITY, "%d ghosts on %d live particles exceed capacity %d "
                "(upload again)", hv[0], n_old, c->cap);
  }
  if (hv[1] || hv[2]) {
    c->loaded = false;
    int rc = check_domain(c);
    return rc ? rc : fail(c, ARM_ERR_DOMAIN, "ghost outside the hash grid");
  }
  c->ng_new = hv[0];
  return ARM_OK;
}

// One step of time_integration.f:131-235, fused:
//   sort -> reorder + predictor -> density -> eos [-> objective] -> forces
//   [-> tape] -> corrector + keys of the next step
struct StepHooks {
  const uint8_t* objmask = nullptr;  // objective: jpart += dt sum f (p^2 or u^2)
  double* jpart = nullptr;
  int kind = ARM_OBJ_PRESSURE;
  const void* tape = nullptr;        // Tape<R>: record the step after the force pass
};

template <typename R> int fused_step(arm_ctx* c, int its, const StepHooks& hk) {
  const bool xs = (c->prm.flags & ARM_AV_VEL) != 0;
  int rc;
  c->its_next = its;
  c->K.its = its;
  const uint32_t* sval = nullptr;
  bool ranges = false;
  // with ghosts: the new set goes after the live slots; the old set carries
  // the extra key (corrector) and drops out of the sort
  int n_sort = c->n;
  if (c->ghosts) {
    if ((rc = make_ghosts<R>(c, true))) return rc;
    n_sort = c->n + c->ng_new;
  }
  {
    PassTimer tm(c, ARM_PASS_SORT);
    if ((rc = run_sort(c, stA<R>(c).id, &sval, &ranges, n_sort))) return rc;
  }
  if (c->ghosts) {
    c->ng = c->ng_new;
    c->n = c->K.n = c->nb + c->ng;
  }
  c->have_ranges = true;
  const int n = c->n;
  {
    St<R>& A = stA<R>(c);
    St<R>& B = stB<R>(c);
    PassTimer tm(c, ARM_PASS_REORDER);
    // with the sort's cell ranges the kernel also adds the row pieces
    const RowPieces rp = ranges ? row_pieces_of(c) : RowPieces{};
    hipLaunchKernelGGL(k_reorder_pred<R>, dim3(grid_of(n) + row_blocks(rp)),
                       dim3(BLOCK), 0, c->stream, A, B, c->K, c->skey, sval, c->cstart,
                       its != 1 ? 1 : 0, ranges ? 0 : 1, c->G, rp);
    c->pieces_valid = rp.pieces != nullptr;
    std::swap(A, B);
    c->masks_valid = false;
  }
  if (!xs && (rc = zero_av<R>(c))) return rc;
  c->sorted = true;
  if ((rc = density_t<R>(c, c->prm.density_method, true))) return rc;
  if (hk.jpart)
    hipLaunchKernelGGL(k_objective<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, stA<R>(c),
                       c->K, hk.kind, hk.objmask, hk.jpart);
  if ((rc = force_t<R>(c, true, c->prm.flags, xs))) return rc;
  take_new_h<R>(c);
  if (hk.tape) {
    const int nb = std::max(grid_of(n), grid_of(c->K.ncell + 1));
    hipLaunchKernelGGL(k_tape<R>, dim3(nb), dim3(BLOCK), 0, c->stream, stA<R>(c), c->K, c->skey,
                       c->cstart, c->masks_valid ? c->nbm : nullptr,
                       *static_cast<const Tape<R>*>(hk.tape));
  }
  if ((rc = corrector_t<R>(c, its, true))) return rc;
  return ARM_OK;
}

template <typename R> int time_integration_t(arm_ctx* c, int its0, int nsteps) {
  int rc;
  // ghosts are selected on a sorted state (cell ranges of the last sort)
  if (c->ghosts && (!c->have_ranges || c->ghost_pending) && (rc = link_list_t<R>(c))) return rc;
  FusedRun run(c, its0);
  if ((rc = begin_steps<R>(c))) return rc;
  if (!(c->prm.flags & ARM_AV_VEL) && (rc = zero_av<R>(c))) return rc;
  const StepHooks none;
  for (int its = its0; its < its0 + nsteps; ++its)
    if ((rc = fused_step<R>(c, its, none))) return rc;
  c->have_rates = true;
  HIPCHK(c, hipGetLastError());
  return check_domain(c, its0 + nsteps);
}

// ---- adjoint rollout (FP64) ---------------------------------------------------
// Device buffers owned by one call.
struct Scratch {
  arm_ctx* c;
  std::vector<void*> ptrs;
  explicit Scratch(arm_ctx* c_) : c(c_) {}
  ~Scratch() {
    for (void* q : ptrs) (void)hipFree(q);
  }
  template <typename T> int get(T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess)
      return fail(c, ARM_ERR_HIP, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
    ptrs.push_back(q);
    *p = static_cast<T*>(q);
    return ARM_OK;
  }
};

// the part of the state a step reads at its start
template <typename R> struct Checkpoint {
  using R2 = typename V2<R>::type;
  int32_t* id;
  double2* xy;
  R2 *v, *a;
  R *rho, *p, *du, *drho;
  double* u;
};

template <typename R> int alloc_checkpoint(Scratch& sc, Checkpoint<R>& k, int n) {
  int rc;
  if ((rc = sc.get(&k.id, n)) || (rc = sc.get(&k.xy, n)) || (rc = sc.get(&k.v, n)) ||
      (rc = sc.get(&k.a, n)) || (rc = sc.get(&k.rho, n)) || (rc = sc.get(&k.u, n)) ||
      (rc = sc.get(&k.p, n)) || (rc = sc.get(&k.du, n)) || (rc = sc.get(&k.drho, n)))
    return rc;
  return ARM_OK;
}

template <typename R>
int copy_checkpoint(arm_ctx* c, const St<R>& S, const Checkpoint<R>& k, bool save) {
  using R2 = typename V2<R>::type;
  const size_t n = (size_t)c->n;
  auto cp = [&](void* st, void* ck, size_t bytes) {
    return save ? hipMemcpyAsync(ck, st, bytes, hipMemcpyDeviceToDevice, c->stream)
                : hipMemcpyAsync(st, ck, bytes, hipMemcpyDeviceToDevice, c->stream);
  };
  HIPCHK(c, cp(S.id, k.id, n * sizeof(int32_t)));
  HIPCHK(c, cp(S.xy, k.xy, n * sizeof(double2)));
  HIPCHK(c, cp(S.v, k.v, n * sizeof(R2)));
  HIPCHK(c, cp(S.a, k.a, n * sizeof(R2)));
  HIPCHK(c, cp(S.rho, k.rho, n * sizeof(R)));
  HIPCHK(c, cp(S.u, k.u, n * sizeof(double)));
  HIPCHK(c, cp(S.p, k.p, n * sizeof(R)));
  HIPCHK(c, cp(S.du, k.du, n * sizeof(R)));
  HIPCHK(c, cp(S.drho, k.drho, n * sizeof(R)));
  return ARM_OK;
}

template <typename R> int alloc_tape(Scratch& sc, Tape<R>& t, int n, int ncell, bool visc) {
  int rc;
  t.cxy = nullptr;
  t.off = nullptr;
  t.tau = nullptr;
  if (sizeof(R) == 4 && ((rc = sc.get(&t.cxy, n)) || (rc = sc.get(&t.off, n)))) return rc;
  if (visc && (rc = sc.get(&t.tau, 4 * (size_t)n))) return rc;
  if ((rc = sc.get(&t.id, n)) || (rc = sc.get(&t.xy, n)) || (rc = sc.get(&t.v, n)) ||
      (rc = sc.get(&t.rho, n)) || (rc = sc.get(&t.u, n)) || (rc = sc.get(&t.psi, n)) ||
      (rc = sc.get(&t.divv, n)) || (rc = sc.get(&t.p, n)) || (rc = sc.get(&t.drho, n)) ||
      (rc = sc.get(&t.du, n)) || (rc = sc.get(&t.a, n)) || (rc = sc.get(&t.av, n)) ||
      (rc = sc.get(&t.rho_min, n)) || (rc = sc.get(&t.u_min, n)) || (rc = sc.get(&t.v_min, n)) ||
      (rc = sc.get(&t.flag, n)) || (rc = sc.get(&t.skey, n)) || (rc = sc.get(&t.cstart, ncell + 1)) ||
      (rc = sc.get(&t.nbm, n)))
    return rc;
  return ARM_OK;
}

template <typename R, bool AV, bool AH, bool XS, bool VI>
void launch_adj_force(arm_ctx* c, const Tape<R>& T, const AdjScratch<R>& W, double* apart,
                      double* qpart, double* epart) {
  hipLaunchKernelGGL((k_adj_force<R, AV, AH, XS, VI>), dim3(grid_of(c->n)), dim3(BLOCK), 0,
                     c->stream, T, c->K, W, apart, qpart, epart);
}

template <typename R, bool VI>
void adj_force_v(arm_ctx* c, const Tape<R>& T, const AdjScratch<R>& W, double* apart,
                 double* qpart, double* epart) {
  const int f = c->prm.flags;
  const int code = ((f & ARM_ART_VISC) ? 4 : 0) | ((f & ARM_ART_HEAT) ? 2 : 0) | ((f & ARM_AV_VEL) ? 1 : 0);
  switch (code) {
    case 0: launch_adj_force<R, false, false, false, VI>(c, T, W, apart, qpart, epart); break;
    case 1: launch_adj_force<R, false, false, true, VI>(c, T, W, apart, qpart, epart); break;
    case 2: launch_adj_force<R, false, true, false, VI>(c, T, W, apart, qpart, epart); break;
    case 3: launch_adj_force<R, false, true, true, VI>(c, T, W, apart, qpart, epart); break;
    case 4: launch_adj_force<R, true, false, false, VI>(c, T, W, apart, qpart, epart); break;
    case 5: launch_adj_force<R, true, false, true, VI>(c, T, W, apart, qpart, epart); break;
    case 6: launch_adj_force<R, true, true, false, VI>(c, T, W, apart, qpart, epart); break;
    default: launch_adj_force<R, true, true, true, VI>(c, T, W, apart, qpart, epart); break;
  }
}

// force-pass adjoint, then (physical viscosity) the shear-tensor adjoint
template <typename R>
void adj_force(arm_ctx* c, const Tape<R>& T, const AdjScratch<R>& W, double* apart,
               double* qpart, double* epart) {
  if (c->prm.flags & ARM_VISC) {
    adj_force_v<R, true>(c, T, W, apart, qpart, epart);
    hipLaunchKernelGGL(k_adj_tensor<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, T, c->K,
                       W);
  } else {
    adj_force_v<R, false>(c, T, W, apart, qpart, epart);
  }
}

}  // namespace

// the part of the 3D state a step reads at its start
template <typename R> struct Checkpoint3 {
  using R4 = typename V4<R>::type;
  int32_t* id;
  double4* xyz;
  R4 *v, *a;
  R *rho, *p, *du, *drho;
  double* u;
};

template <typename R> struct AdjArrays3 {
  std::vector<Checkpoint3<R>> ck;
  std::vector<Tape3<R>> tape;
  AdjScratch3<R> W{};
};

// Checkpoints, tapes and adjoint arrays of one rollout, kept between
// arm_rollout_forward and arm_rollout_backward.
template <typename R> struct AdjArrays {
  std::vector<Checkpoint<R>> ck;
  std::vector<Tape<R>> tape;
  AdjScratch<R> W{};
};

struct AdjRun {
  Scratch sc;
  int nsteps = 0, ckpt = 0, nck = 0;
  uint8_t* obj = nullptr;
  double *jpart = nullptr, *apart = nullptr, *qpart = nullptr, *epart = nullptr;
  int kind = ARM_OBJ_PRESSURE;
  AdjState L{};
  AdjArrays<double> a64;
  AdjArrays<float> a32;
  template <typename R> AdjArrays<R>& arrays() {
    if constexpr (sizeof(R) == 8)
      return a64;
    else
      return a32;
  }
  AdjState3 L3{};
  AdjArrays3<double> b64;
  AdjArrays3<float> b32;
  template <typename R> AdjArrays3<R>& arrays3() {
    if constexpr (sizeof(R) == 8)
      return b64;
    else
      return b32;
  }
  explicit AdjRun(arm_ctx* c) : sc(c) {}
};

namespace {

void drop_adjoint(arm_ctx* c) {
  delete c->adj;
  c->adj = nullptr;
}

template <typename Rf>
int rollout_forward(arm_ctx* c, int nsteps, int ckpt, const int32_t* obj_ids, int nobj, double* J,
                    double* ms) {
  const int n = c->n, ncell = c->K.ncell;
  const int nb = grid_of(n);
  drop_adjoint(c);
  const bool visc = (c->prm.flags & ARM_VISC) != 0;
  AdjRun* R = new AdjRun(c);
  c->adj = R;
  Scratch& sc = R->sc;
  int rc;
  R->nsteps = nsteps;
  R->ckpt = ckpt;
  R->nck = (nsteps + ckpt - 1) / ckpt;
  R->kind = c->objkind;
  if ((rc = sc.get(&R->obj, n)) || (rc = sc.get(&R->jpart, nb)) || (rc = sc.get(&R->apart, nb)) ||
      (rc = sc.get(&R->qpart, nb)) || (rc = sc.get(&R->epart, nb)))
    return rc;
  {
    std::vector<uint8_t> h(n, 0);
    for (int k = 0; k < nobj; ++k) {
      if (obj_ids[k] < 0 || obj_ids[k] >= n)
        return fail(c, ARM_ERR_ARG, "objective id %d out of range", obj_ids[k]);
      h[obj_ids[k]] = 1;
    }
    HIPCHK(c, hipMemcpyAsync(R->obj, h.data(), n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  HIPCHK(c, hipMemsetAsync(R->jpart, 0, sizeof(double) * nb, c->stream));
  AdjArrays<Rf>& Ar = R->arrays<Rf>();
  Ar.ck.resize(R->nck + 1);
  for (auto& k : Ar.ck)
    if ((rc = alloc_checkpoint(sc, k, n))) return rc;
  Ar.tape.resize(ckpt);
  for (auto& t : Ar.tape)
    if ((rc = alloc_tape(sc, t, n, ncell, visc))) return rc;
  double** lf[] = {&R->L.x,   &R->L.y,   &R->L.vx,  &R->L.vy,  &R->L.rho,   &R->L.u,
                   &R->L.bx,  &R->L.by,  &R->L.bvx, &R->L.bvy, &R->L.brho,  &R->L.bu,
                   &R->L.cax, &R->L.cay, &R->L.cdrho, &R->L.cdu};
  for (double** q : lf)
    if ((rc = sc.get(q, n))) return rc;
  AdjScratch<Rf>& W = Ar.W;
  Rf** wf[] = {&W.lax, &W.lay, &W.ldrho, &W.ldu, &W.lavx, &W.lavy, &W.lp,  &W.lQ,  &W.lA,
               &W.lpsi, &W.lS, &W.lrhoh, &W.lu,  &W.lvx,  &W.lvy,  &W.lx,  &W.ly,  &W.lus};
  for (Rf** q : wf)
    if ((rc = sc.get(q, n))) return rc;
  W.lTxx = W.lTyy = W.lTxy = nullptr;
  if (visc && ((rc = sc.get(&W.lTxx, n)) || (rc = sc.get(&W.lTyy, n)) || (rc = sc.get(&W.lTxy, n))))
    return rc;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  if ((rc = begin_steps<Rf>(c))) return rc;
  if (!(c->prm.flags & ARM_AV_VEL) && (rc = zero_av<Rf>(c))) return rc;
  StepHooks fwd;
  fwd.objmask = R->obj;
  fwd.jpart = R->jpart;
  fwd.kind = R->kind;
  for (int its = 1; its <= nsteps; ++its) {
    if ((its - 1) % ckpt == 0)
      if ((rc = copy_checkpoint(c, stA<Rf>(c), Ar.ck[(its - 1) / ckpt], true))) return rc;
    if ((rc = fused_step<Rf>(c, its, fwd))) return rc;
  }
  if ((rc = copy_checkpoint(c, stA<Rf>(c), Ar.ck[R->nck], true))) return rc;  // final state
  HIPCHK(c, hipEventRecord(e1, c->stream));
  std::vector<double> jp(nb);
  HIPCHK(c, hipMemcpyAsync(jp.data(), R->jpart, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float f = 0.f;
  (void)hipEventElapsedTime(&f, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  double Js = 0;
  for (int k = 0; k < nb; ++k) Js += jp[k];
  if (J) *J = Js;
  if (ms) *ms = f;
  c->sorted = false;
  c->masks_valid = false;
  c->have_rates = true;
  return check_domain(c);
}

template <typename Rf>
int rollout_backward(arm_ctx* c, double seed, double* dJ_dalpha, double* dJ_dv0, double* ms) {
  AdjRun* R = c->adj;
  const int n = c->n, nb = grid_of(n);
  const int nsteps = R->nsteps, ckpt = R->ckpt, nck = R->nck;
  AdjState& L = R->L;
  AdjArrays<Rf>& Ar = R->arrays<Rf>();
  AdjScratch<Rf>& W = Ar.W;
  int rc;
  double** lf[] = {&L.x, &L.y, &L.vx, &L.vy, &L.rho, &L.u, &L.cax, &L.cay, &L.cdrho, &L.cdu};
  for (double** q : lf) HIPCHK(c, hipMemsetAsync(*q, 0, sizeof(double) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(R->apart, 0, sizeof(double) * nb, c->stream));
  HIPCHK(c, hipMemsetAsync(R->qpart, 0, sizeof(double) * nb, c->stream));
  HIPCHK(c, hipMemsetAsync(R->epart, 0, sizeof(double) * nb, c->stream));
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  for (int seg = nck - 1; seg >= 0; --seg) {
    const int s0 = 1 + seg * ckpt, s1 = std::min(s0 + ckpt, nsteps + 1);
    if ((rc = copy_checkpoint(c, stA<Rf>(c), Ar.ck[seg], false))) return rc;
    if ((rc = begin_steps<Rf>(c))) return rc;
    for (int its = s0; its < s1; ++its) {
      StepHooks rec;
      rec.tape = &Ar.tape[its - s0];
      if ((rc = fused_step<Rf>(c, its, rec))) return rc;
    }
    for (int its = s1 - 1; its >= s0; --its) {
      const Tape<Rf>& T = Ar.tape[its - s0];
      hipLaunchKernelGGL(k_adj_corrector<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, its,
                         L, W, R->obj, R->kind, seed);
      adj_force<Rf>(c, T, W, R->apart, R->qpart, R->epart);
      hipLaunchKernelGGL(k_adj_eos<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, W);
      hipLaunchKernelGGL(k_adj_density<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, W);
      hipLaunchKernelGGL(k_adj_predictor<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, its,
                         L, W);
      HIPCHK(c, hipGetLastError());
      // the state adjoint before this step is the one after the previous step
      std::swap(L.x, L.bx);
      std::swap(L.y, L.by);
      std::swap(L.vx, L.bvx);
      std::swap(L.vy, L.bvy);
      std::swap(L.rho, L.brho);
      std::swap(L.u, L.bu);
    }
  }
  HIPCHK(c, hipEventRecord(e1, c->stream));
  // leave the context at the final forward state
  if ((rc = copy_checkpoint(c, stA<Rf>(c), Ar.ck[nck], false))) return rc;
  std::vector<double> ap(nb), qp(nb), ep(nb), gx(n), gy(n);
  HIPCHK(c, hipMemcpyAsync(ap.data(), R->apart, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(qp.data(), R->qpart, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(ep.data(), R->epart, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(gx.data(), L.vx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(gy.data(), L.vy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float f = 0.f;
  (void)hipEventElapsedTime(&f, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  double As = 0, Qs = 0, Es = 0;
  for (int k = 0; k < nb; ++k) {
    As += ap[k];
    Qs += qp[k];
    Es += ep[k];
  }
  c->grads[ARM_GRAD_ALPHA] = As;
  c->grads[ARM_GRAD_Q0] = Qs;
  c->grads[ARM_GRAD_VISC] = Es;
  if (dJ_dalpha) *dJ_dalpha = As;
  if (dJ_dv0)
    for (int i = 0; i < n; ++i) {
      dJ_dv0[2 * i] = gx[i];
      dJ_dv0[2 * i + 1] = gy[i];
    }
  if (ms) *ms = f;
  c->sorted = false;
  c->masks_valid = false;
  return check_domain(c);
}

template <typename R> int pairs_t(arm_ctx* c, int cap, int* niac, int* pi, int* pj) {
  St<R>& A = stA<R>(c);
  const int n = c->n;
  hipLaunchKernelGGL(k_pair_count<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                     c->skey, c->cstart, c->d_cnt, nullptr, nullptr, nullptr);
  HIPCHK(c, hipGetLastError());
  std::vector<int> cnt(n);
  HIPCHK(c, hipMemcpyAsync(cnt.data(), c->d_cnt, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  long long total = 0;
  std::vector<int> offs(n);
  for (int i = 0; i < n; ++i) {
    offs[i] = (int)total;
    total += cnt[i];
  }
  if (niac) *niac = (int)total;
  if (!pi) return ARM_OK;
  if (total > cap) return fail(c, ARM_ERR_CAPACITY, "%lld pairs exceed cap %d", total, cap);
  int *d_off = nullptr, *d_i = nullptr, *d_j = nullptr;
  HIPCHK(c, hipMalloc(&d_off, sizeof(int) * std::max(n, 1)));
  HIPCHK(c, hipMalloc(&d_i, sizeof(int) * std::max<long long>(total, 1)));
  HIPCHK(c, hipMalloc(&d_j, sizeof(int) * std::max<long long>(total, 1)));
  (void)hipMemcpyAsync(d_off, offs.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream);
  hipLaunchKernelGGL(k_pair_count<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                     c->skey, c->cstart, nullptr, d_off, d_i, d_j);
  std::vector<int> vi(total), vj(total);
  (void)hipMemcpyAsync(vi.data(), d_i, sizeof(int) * total, hipMemcpyDeviceToHost, c->stream);
  (void)hipMemcpyAsync(vj.data(), d_j, sizeof(int) * total, hipMemcpyDeviceToHost, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_off);
  (void)hipFree(d_i);
  (void)hipFree(d_j);
  if (e != hipSuccess) return fail(c, ARM_ERR_HIP, "pair emission failed: %s", hipGetErrorString(e));
  std::vector<unsigned long long> kk(total);
  for (long long k = 0; k < total; ++k)
    kk[k] = ((unsigned long long)(unsigned)vi[k] << 32) | (unsigned)vj[k];
  std::sort(kk.begin(), kk.end());
  for (long long k = 0; k < total; ++k) {
    pi[k] = (int)(kk[k] >> 32);
    pj[k] = (int)(kk[k] & 0xffffffffu);
  }
  return ARM_OK;
}

// ---- host <-> device --------------------------------------------------------
template <typename R> int upload_t(arm_ctx* c, const double* x, const double* v, const double* rho,
                                   const double* p, const double* u) {
  using R2 = typename V2<R>::type;
  const int n = c->n;
  St<R>& A = stA<R>(c);
  std::vector<int32_t> id(n);
  std::vector<double2> xy(n);
  std::vector<R2> vv(n);
  std::vector<R> rr(n), pp(n);
  std::vector<double> uu(n);
  std::vector<float> uf(n);
  for (int i = 0; i < n; ++i) {
    id[i] = i;
    xy[i] = make_double2(x[2 * i], x[2 * i + 1]);
    vv[i].x = (R)(v ? v[2 * i] : 0.0);
    vv[i].y = (R)(v ? v[2 * i + 1] : 0.0);
    const double r = rho ? rho[i] : c->prm.rho0;
    rr[i] = Layout<R>::f32 ? (R)(r - c->prm.rho0) : (R)r;
    pp[i] = (R)(p ? p[i] : 0.0);
    uu[i] = u ? u[i] : 0.0;
    uf[i] = (float)uu[i];
  }
  auto cp = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  };
  HIPCHK(c, cp(A.id, id.data(), sizeof(int32_t) * n));
  HIPCHK(c, cp(A.xy, xy.data(), sizeof(double2) * n));
  HIPCHK(c, cp(A.v, vv.data(), sizeof(R2) * n));
  HIPCHK(c, cp(A.rho, rr.data(), sizeof(R) * n));
  HIPCHK(c, cp(A.p, pp.data(), sizeof(R) * n));
  HIPCHK(c, cp(A.u, uu.data(), sizeof(double) * n));
  if (Layout<R>::f32) HIPCHK(c, cp(A.uf, uf.data(), sizeof(float) * n));
  // rates start at zero
  HIPCHK(c, hipMemsetAsync(A.drho, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.du, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.a, 0, sizeof(R2) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.av, 0, sizeof(R2) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.divv, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.psi, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.rho_min, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.u_min, 0, sizeof(double) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.v_min, 0, sizeof(R2) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.flag, 0, n, c->stream));
  {
    std::vector<double> hh(n, c->prm.h);
    HIPCHK(c, cp(A.hs, hh.data(), sizeof(double) * n));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  // FP32 form: cell + offset from the FP64 positions (the keys kernel writes them)
  if (Layout<R>::f32) {
    hipLaunchKernelGGL(k_keys<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                       KeyOut{}, c->d_err, 0);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_domain(c);
}

template <typename T> int d2h(arm_ctx* c, std::vector<T>& dst, const T* src, int n) {
  dst.resize(n);
  HIPCHK(c, hipMemcpyAsync(dst.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  return ARM_OK;
}

template <typename R> int download_t(arm_ctx* c, double* x, double* v, double* rho, double* p,
                                     double* u) {
  using R2 = typename V2<R>::type;
  const int n = c->n;
  St<R>& A = stA<R>(c);
  std::vector<int32_t> id;
  std::vector<double2> xy;
  std::vector<R2> vv;
  std::vector<R> rr, pp;
  std::vector<double> uu;
  int rc;
  if ((rc = d2h(c, id, A.id, n))) return rc;
  if (x && (rc = d2h(c, xy, A.xy, n))) return rc;
  if (v && (rc = d2h(c, vv, A.v, n))) return rc;
  if (rho && (rc = d2h(c, rr, A.rho, n))) return rc;
  if (p && (rc = d2h(c, pp, A.p, n))) return rc;
  if (u && (rc = d2h(c, uu, A.u, n))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < n; ++s) {
    const int i = id[s];
    if (i >= c->nb) continue;  // ghosts: arm_download_ghosts
    if (x) {
      x[2 * i] = xy[s].x;
      x[2 * i + 1] = xy[s].y;
    }
    if (v) {
      v[2 * i] = (double)vv[s].x;
      v[2 * i + 1] = (double)vv[s].y;
    }
    if (rho) rho[i] = Layout<R>::f32 ? c->prm.rho0 + (double)rr[s] : (double)rr[s];
    if (p) p[i] = (double)pp[s];
    if (u) u[i] = uu[s];
  }
  return ARM_OK;
}

template <typename R> int download_ghosts_t(arm_ctx* c, double* x, double* v, double* rho,
                                            double* p, double* u, int32_t* parent) {
  using R2 = typename V2<R>::type;
  const int n = c->n;
  St<R>& A = stA<R>(c);
  std::vector<int32_t> id, pid;
  std::vector<double2> xy;
  std::vector<R2> vv;
  std::vector<R> rr, pp;
  std::vector<double> uu;
  int rc;
  if ((rc = d2h(c, id, A.id, n))) return rc;
  if ((rc = d2h(c, pid, c->G.gpid, n))) return rc;
  if (x && (rc = d2h(c, xy, A.xy, n))) return rc;
  if (v && (rc = d2h(c, vv, A.v, n))) return rc;
  if (rho && (rc = d2h(c, rr, A.rho, n))) return rc;
  if (p && (rc = d2h(c, pp, A.p, n))) return rc;
  if (u && (rc = d2h(c, uu, A.u, n))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < n; ++s) {
    const int g = id[s] - c->nb;
    if (g < 0) continue;
    if (x) {
      x[2 * g] = xy[s].x;
      x[2 * g + 1] = xy[s].y;
    }
    if (v) {
      v[2 * g] = (double)vv[s].x;
      v[2 * g + 1] = (double)vv[s].y;
    }
    if (rho) rho[g] = Layout<R>::f32 ? c->prm.rho0 + (double)rr[s] : (double)rr[s];
    if (p) p[g] = (double)pp[s];
    if (u) u[g] = uu[s];
    if (parent) parent[g] = pid[s];
  }
  return ARM_OK;
}

template <typename R> int download_rates_t(arm_ctx* c, double* drho, double* dvdt, double* dudt,
                                           double* av, double* divv) {
  using R2 = typename V2<R>::type;
  const int n = c->n;
  St<R>& A = stA<R>(c);
  std::vector<int32_t> id;
  std::vector<R> dr, du, dv;
  std::vector<R2> aa, ww;
  int rc;
  if ((rc = d2h(c, id, A.id, n))) return rc;
  if (drho && (rc = d2h(c, dr, A.drho, n))) return rc;
  if (dudt && (rc = d2h(c, du, A.du, n))) return rc;
  if (divv && (rc = d2h(c, dv, A.divv, n))) return rc;
  if (dvdt && (rc = d2h(c, aa, A.a, n))) return rc;
  if (av && (rc = d2h(c, ww, A.av, n))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < n; ++s) {
    const int i = id[s];
    if (i >= c->nb) continue;  // ghosts carry no rates of their own
    if (drho) drho[i] = (double)dr[s];
    if (dudt) dudt[i] = (double)du[s];
    if (divv) divv[i] = (double)dv[s];
    if (dvdt) {
      dvdt[2 * i] = (double)aa[s].x;
      dvdt[2 * i + 1] = (double)aa[s].y;
    }
    if (av) {
      av[2 * i] = (double)ww[s].x;
      av[2 * i + 1] = (double)ww[s].y;
    }
  }
  return ARM_OK;
}

template <typename R> int conservation_t(arm_ctx* c, double out[5]) {
  St<R>& A = stA<R>(c);
  const int nb = grid_of(c->n);
  hipLaunchKernelGGL(k_conserve<R>, dim3(nb), dim3(BLOCK), 0, c->stream, A, c->K, c->d_part);
  HIPCHK(c, hipGetLastError());
  std::vector<double> part((size_t)nb * 5);
  HIPCHK(c, hipMemcpyAsync(part.data(), c->d_part, sizeof(double) * part.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 5; ++k) out[k] = 0.0;
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 5; ++k) out[k] += part[(size_t)b * 5 + k];
  return ARM_OK;
}

// ---- three dimensions ---------------------------------------------------------
// The 3D step on the same context machinery (sort workspace, cell ranges,
// timing); only the state layout and the gathers differ.
template <typename R> int alloc_state3(arm_ctx* c, St3<R>& S, int cap) {
  int rc = ARM_OK;
#define AL(f, cnt) \
  if ((rc = dalloc(c, &S.f, (cnt))) != ARM_OK) return rc;
  AL(id, cap);
  AL(xyz, cap);
  if (sizeof(R) == 4) {
    AL(pf, (size_t)cap + 8);  // the eight-wide neighbour pre-test reads up to 7 past a row
    AL(uf, cap);
  }
  AL(v, cap);
  AL(rho, cap);
  AL(u, cap);
  AL(p, cap);
  AL(psi, cap);
  AL(drho, cap);
  AL(du, cap);
  AL(a, cap);
  AL(av, cap);
  AL(divv, cap);
  AL(rho_min, cap);
  AL(u_min, cap);
  AL(v_min, cap);
  AL(flag, cap);
#undef AL
  return ARM_OK;
}

bool is3d(const arm_ctx* c) { return c->prm.dim == 3; }

template <typename R> St3<R>& st3A(arm_ctx* c);
template <> St3<double>& st3A<double>(arm_ctx* c) { return c->A3_64; }
template <> St3<float>& st3A<float>(arm_ctx* c) { return c->A3_32; }
template <typename R> St3<R>& st3B(arm_ctx* c);
template <> St3<double>& st3B<double>(arm_ctx* c) { return c->B3_64; }
template <> St3<float>& st3B<float>(arm_ctx* c) { return c->B3_32; }

template <typename R> int link_list3_t(arm_ctx* c) {
  const int n = c->n;
  St3<R>& A = st3A<R>(c);
  St3<R>& B = st3B<R>(c);
  {
    PassTimer tm(c, ARM_PASS_KEYS);
    zero_counts(c);
    hipLaunchKernelGGL(k3_keys<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                       key_out(c), c->d_err);
  }
  const uint32_t* sval = nullptr;
  bool ranges = false;
  {
    PassTimer tm(c, ARM_PASS_SORT);
    int rc = run_sort(c, A.id, &sval, &ranges);
    if (rc) return rc;
  }
  {
    PassTimer tm(c, ARM_PASS_REORDER);
    hipLaunchKernelGGL(k3_reorder_all<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, B,
                       c->K, c->skey, sval, c->cstart, ranges ? 0 : 1);
  }
  HIPCHK(c, hipGetLastError());
  std::swap(A, B);
  c->sorted = true;
  c->masks_valid = false;
  return check_domain(c);
}

// 3D density gather (direct: per-particle walks of the nine rows; an
// LDS-tiled variant measured twice as slow, the 27-cell halo filling the
// LDS for few lanes).  Records the neighbour masks for the force pass.
template <typename R, bool SUM, bool EOS> void launch_density3(arm_ctx* c) {
  hipLaunchKernelGGL((k3_density<R, SUM, EOS>), dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                     st3A<R>(c), c->K, c->skey, c->cstart, Mask3{c->nbm3, c->cap});
  c->masks_valid = true;
}

template <typename R> int density3_t(arm_ctx* c, int method, bool eos) {
  PassTimer tm(c, ARM_PASS_DENSITY);
  if (method == 1)
    eos ? launch_density3<R, true, true>(c) : launch_density3<R, true, false>(c);
  else
    eos ? launch_density3<R, false, true>(c) : launch_density3<R, false, false>(c);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

template <typename R> int eos3_t(arm_ctx* c) {
  PassTimer tm(c, ARM_PASS_EOS);
  hipLaunchKernelGGL(k3_eos<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, st3A<R>(c), c->K);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

template <typename R> Pack3<R> pack3(arm_ctx* c) {
  using R4 = typename V4<R>::type;
  return Pack3<R>{static_cast<R4*>(c->pack3_vq)};
}

template <typename R, bool P, bool AV, bool AH, bool XS> void launch_force3(arm_ctx* c) {
  const Mask3 m{c->masks_valid ? c->nbm3 : nullptr, c->cap};
  const Pack3<R> pk = pack3<R>(c);
  hipLaunchKernelGGL(k3_force_pack<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                     st3A<R>(c), c->K, pk);
  hipLaunchKernelGGL((k3_force<R, P, AV, AH, XS>), dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                     st3A<R>(c), pk, c->K, c->skey, c->cstart, m);
}

template <typename R> int force3_t(arm_ctx* c, bool press, int flags, bool xs) {
  const bool av = (flags & ARM_ART_VISC) != 0, ah = (flags & ARM_ART_HEAT) != 0;
  if (press && (flags & ARM_VISC))
    return fail(c, ARM_ERR_UNSUPPORTED, "physical viscosity is implemented in 2D only");
  dep_window(c);
  PassTimer tm(c, ARM_PASS_FORCE);
  const int code = (press ? 8 : 0) | (av ? 4 : 0) | (ah ? 2 : 0) | (xs ? 1 : 0);
  switch (code) {
    case 8: launch_force3<R, true, false, false, false>(c); break;
    case 9: launch_force3<R, true, false, false, true>(c); break;
    case 10: launch_force3<R, true, false, true, false>(c); break;
    case 11: launch_force3<R, true, false, true, true>(c); break;
    case 12: launch_force3<R, true, true, false, false>(c); break;
    case 13: launch_force3<R, true, true, false, true>(c); break;
    case 14: launch_force3<R, true, true, true, false>(c); break;
    case 15: launch_force3<R, true, true, true, true>(c); break;
    case 1: launch_force3<R, false, false, false, true>(c); break;
    default: return fail(c, ARM_ERR_ARG, "force pass with nothing to compute");
  }
  if (press && (c->K.flags & KF_DEPOSITION))
    hipLaunchKernelGGL(k3_deposition<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                       st3A<R>(c), c->K);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

template <typename R> int zero_av3(arm_ctx* c) {
  return zero_bytes(c, st3A<R>(c).av, sizeof(typename V4<R>::type) * (size_t)c->n);
}

template <typename R> int single_step3_t(arm_ctx* c) {
  int rc;
  if ((rc = density3_t<R>(c, c->prm.density_method, true))) return rc;
  const bool xs = (c->prm.flags & ARM_AV_VEL) != 0;
  if ((rc = force3_t<R>(c, true, c->prm.flags, xs))) return rc;
  if (!xs && (rc = zero_av3<R>(c))) return rc;
  c->have_rates = true;
  return ARM_OK;
}

template <typename R> int predictor3_t(arm_ctx* c) {
  PassTimer tm(c, ARM_PASS_INTEGRATE);
  hipLaunchKernelGGL(k3_predictor<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                     st3A<R>(c), c->K);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

template <typename R> int corrector3_t(arm_ctx* c, int its, bool keys) {
  c->its_next = its + 1;
  PassTimer tm(c, ARM_PASS_INTEGRATE);
  hipLaunchKernelGGL(k3_corrector<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                     st3A<R>(c), c->K, its, keys ? 1 : 0, keys ? key_out(c) : KeyOut{},
                     c->d_err);
  HIPCHK(c, hipGetLastError());
  return ARM_OK;
}

// The fused step of time_integration.f:131-235 in 3D, as fused_step in 2D.
template <typename R> int fused_step3(arm_ctx* c, int its, const StepHooks& hk = StepHooks{}) {
  const int n = c->n;
  const bool xs = (c->prm.flags & ARM_AV_VEL) != 0;
  int rc;
  c->its_next = its;
  c->K.its = its;
  const uint32_t* sval = nullptr;
  bool ranges = false;
  {
    PassTimer tm(c, ARM_PASS_SORT);
    if ((rc = run_sort(c, st3A<R>(c).id, &sval, &ranges))) return rc;
  }
  {
    St3<R>& A = st3A<R>(c);
    St3<R>& B = st3B<R>(c);
    PassTimer tm(c, ARM_PASS_REORDER);
    hipLaunchKernelGGL(k3_reorder_pred<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, B,
                       c->K, c->skey, sval, c->cstart, its != 1 ? 1 : 0, ranges ? 0 : 1);
    std::swap(A, B);
    c->masks_valid = false;
  }
  if (!xs && (rc = zero_av3<R>(c))) return rc;
  c->sorted = true;
  if ((rc = density3_t<R>(c, c->prm.density_method, true))) return rc;
  if (hk.jpart)
    hipLaunchKernelGGL(k3_objective<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, st3A<R>(c),
                       c->K, hk.kind, hk.objmask, hk.jpart);
  if ((rc = force3_t<R>(c, true, c->prm.flags, xs))) return rc;
  if (hk.tape) {
    const int nb = std::max(grid_of(n), grid_of(c->K.ncell + 1));
    hipLaunchKernelGGL(k3_tape<R>, dim3(nb), dim3(BLOCK), 0, c->stream, st3A<R>(c), c->K, c->skey,
                       c->cstart, Mask3{c->masks_valid ? c->nbm3 : nullptr, c->cap},
                       *static_cast<const Tape3<R>*>(hk.tape));
  }
  return corrector3_t<R>(c, its, true);
}

template <typename R> int time_integration3_t(arm_ctx* c, int its0, int nsteps) {
  int rc;
  FusedRun run(c, its0);
  {
    PassTimer tm(c, ARM_PASS_KEYS);
    zero_counts(c);
    hipLaunchKernelGGL(k3_keys<R>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, st3A<R>(c),
                       c->K, key_out(c), c->d_err);
    HIPCHK(c, hipGetLastError());
  }
  if (!(c->prm.flags & ARM_AV_VEL) && (rc = zero_av3<R>(c))) return rc;
  for (int its = its0; its < its0 + nsteps; ++its)
    if ((rc = fused_step3<R>(c, its))) return rc;
  c->have_rates = true;
  HIPCHK(c, hipGetLastError());
  return check_domain(c, its0 + nsteps);
}

template <typename R> int pairs3_t(arm_ctx* c, int cap, int* niac, int* pi, int* pj) {
  St3<R>& A = st3A<R>(c);
  const int n = c->n;
  hipLaunchKernelGGL(k3_pair_count<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                     c->skey, c->cstart, c->d_cnt, nullptr, nullptr, nullptr);
  HIPCHK(c, hipGetLastError());
  std::vector<int> cnt(n);
  HIPCHK(c, hipMemcpyAsync(cnt.data(), c->d_cnt, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  long long total = 0;
  std::vector<int> offs(n);
  for (int i = 0; i < n; ++i) {
    offs[i] = (int)total;
    total += cnt[i];
  }
  if (total > 0x7fffffffLL) return fail(c, ARM_ERR_CAPACITY, "%lld pairs exceed int range", total);
  if (niac) *niac = (int)total;
  if (!pi) return ARM_OK;
  if (total > cap) return fail(c, ARM_ERR_CAPACITY, "%lld pairs exceed cap %d", total, cap);
  int *d_off = nullptr, *d_i = nullptr, *d_j = nullptr;
  HIPCHK(c, hipMalloc(&d_off, sizeof(int) * std::max(n, 1)));
  HIPCHK(c, hipMalloc(&d_i, sizeof(int) * std::max<long long>(total, 1)));
  HIPCHK(c, hipMalloc(&d_j, sizeof(int) * std::max<long long>(total, 1)));
  (void)hipMemcpyAsync(d_off, offs.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream);
  hipLaunchKernelGGL(k3_pair_count<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                     c->skey, c->cstart, nullptr, d_off, d_i, d_j);
  std::vector<int> vi(total), vj(total);
  (void)hipMemcpyAsync(vi.data(), d_i, sizeof(int) * total, hipMemcpyDeviceToHost, c->stream);
  (void)hipMemcpyAsync(vj.data(), d_j, sizeof(int) * total, hipMemcpyDeviceToHost, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_off);
  (void)hipFree(d_i);
  (void)hipFree(d_j);
  if (e != hipSuccess) return fail(c, ARM_ERR_HIP, "pair emission failed: %s", hipGetErrorString(e));
  std::vector<unsigned long long> kk(total);
  for (long long k = 0; k < total; ++k)
    kk[k] = ((unsigned long long)(unsigned)vi[k] << 32) | (unsigned)vj[k];
  std::sort(kk.begin(), kk.end());
  for (long long k = 0; k < total; ++k) {
    pi[k] = (int)(kk[k] >> 32);
    pj[k] = (int)(kk[k] & 0xffffffffu);
  }
  return ARM_OK;
}

template <typename R> int upload3_t(arm_ctx* c, const double* x, const double* v,
                                    const double* rho, const double* u) {
  using R4 = typename V4<R>::type;
  const int n = c->n;
  St3<R>& A = st3A<R>(c);
  std::vector<int32_t> id(n);
  std::vector<double4> xyz(n);
  std::vector<R4> vv(n);
  std::vector<R> rr(n), pp(n, (R)0);
  std::vector<double> uu(n);
  std::vector<float> uf(n);
  for (int i = 0; i < n; ++i) {
    id[i] = i;
    xyz[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0);
    vv[i].x = (R)(v ? v[3 * i] : 0.0);
    vv[i].y = (R)(v ? v[3 * i + 1] : 0.0);
    vv[i].z = (R)(v ? v[3 * i + 2] : 0.0);
    vv[i].w = (R)0;
    const double r = rho ? rho[i] : c->prm.rho0;
    rr[i] = Layout<R>::f32 ? (R)(r - c->prm.rho0) : (R)r;
    uu[i] = u ? u[i] : 0.0;
    uf[i] = (float)uu[i];
  }
  auto cp = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  };
  HIPCHK(c, cp(A.id, id.data(), sizeof(int32_t) * n));
  HIPCHK(c, cp(A.xyz, xyz.data(), sizeof(double4) * n));
  HIPCHK(c, cp(A.v, vv.data(), sizeof(R4) * n));
  HIPCHK(c, cp(A.rho, rr.data(), sizeof(R) * n));
  HIPCHK(c, cp(A.p, pp.data(), sizeof(R) * n));
  HIPCHK(c, cp(A.u, uu.data(), sizeof(double) * n));
  if (Layout<R>::f32) HIPCHK(c, cp(A.uf, uf.data(), sizeof(float) * n));
  HIPCHK(c, hipMemsetAsync(A.drho, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.du, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.a, 0, sizeof(R4) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.av, 0, sizeof(R4) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.divv, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.psi, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.rho_min, 0, sizeof(R) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.u_min, 0, sizeof(double) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.v_min, 0, sizeof(R4) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(A.flag, 0, n, c->stream));
  if (Layout<R>::f32) {
    hipLaunchKernelGGL(k3_keys<R>, dim3(grid_of(n)), dim3(BLOCK), 0, c->stream, A, c->K,
                       KeyOut{}, c->d_err);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_domain(c);
}

template <typename R> int download3_t(arm_ctx* c, double* x, double* v, double* rho, double* p,
                                      double* u) {
  using R4 = typename V4<R>::type;
  const int n = c->n;
  St3<R>& A = st3A<R>(c);
  std::vector<int32_t> id;
  std::vector<double4> xyz;
  std::vector<R4> vv;
  std::vector<R> rr, pp;
  std::vector<double> uu;
  int rc;
  if ((rc = d2h(c, id, A.id, n))) return rc;
  if (x && (rc = d2h(c, xyz, A.xyz, n))) return rc;
  if (v && (rc = d2h(c, vv, A.v, n))) return rc;
  if (rho && (rc = d2h(c, rr, A.rho, n))) return rc;
  if (p && (rc = d2h(c, pp, A.p, n))) return rc;
  if (u && (rc = d2h(c, uu, A.u, n))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < n; ++s) {
    const int i = id[s];
    if (x) {
      x[3 * i] = xyz[s].x;
      x[3 * i + 1] = xyz[s].y;
      x[3 * i + 2] = xyz[s].z;
    }
    if (v) {
      v[3 * i] = (double)vv[s].x;
      v[3 * i + 1] = (double)vv[s].y;
      v[3 * i + 2] = (double)vv[s].z;
    }
    if (rho) rho[i] = Layout<R>::f32 ? c->prm.rho0 + (double)rr[s] : (double)rr[s];
    if (p) p[i] = (double)pp[s];
    if (u) u[i] = uu[s];
  }
  return ARM_OK;
}

template <typename R> int download_rates3_t(arm_ctx* c, double* drho, double* dvdt, double* dudt,
                                            double* av, double* divv) {
  using R4 = typename V4<R>::type;
  const int n = c->n;
  St3<R>& A = st3A<R>(c);
  std::vector<int32_t> id;
  std::vector<R> dr, du, dv;
  std::vector<R4> aa, ww;
  int rc;
  if ((rc = d2h(c, id, A.id, n))) return rc;
  if (drho && (rc = d2h(c, dr, A.drho, n))) return rc;
  if (dudt && (rc = d2h(c, du, A.du, n))) return rc;
  if (divv && (rc = d2h(c, dv, A.divv, n))) return rc;
  if (dvdt && (rc = d2h(c, aa, A.a, n))) return rc;
  if (av && (rc = d2h(c, ww, A.av, n))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < n; ++s) {
    const int i = id[s];
    if (i >= c->nb) continue;  // ghosts carry no rates of their own
    if (drho) drho[i] = (double)dr[s];
    if (dudt) dudt[i] = (double)du[s];
    if (divv) divv[i] = (double)dv[s];
    if (dvdt) {
      dvdt[3 * i] = (double)aa[s].x;
      dvdt[3 * i + 1] = (double)aa[s].y;
      dvdt[3 * i + 2] = (double)aa[s].z;
    }
    if (av) {
      av[3 * i] = (double)ww[s].x;
      av[3 * i + 1] = (double)ww[s].y;
      av[3 * i + 2] = (double)ww[s].z;
    }
  }
  return ARM_OK;
}

template <typename R> int conservation3_t(arm_ctx* c, double out[6]) {
  const int nb = grid_of(c->n);
  hipLaunchKernelGGL(k3_conserve<R>, dim3(nb), dim3(BLOCK), 0, c->stream, st3A<R>(c), c->K,
                     c->d_part);
  HIPCHK(c, hipGetLastError());
  std::vector<double> part((size_t)nb * 6);
  HIPCHK(c, hipMemcpyAsync(part.data(), c->d_part, sizeof(double) * part.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 6; ++k) out[k] = 0.0;
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 6; ++k) out[k] += part[(size_t)b * 6 + k];
  return ARM_OK;
}

// ---- 3D adjoint rollout ------------------------------------------------------
template <typename R> int alloc_checkpoint3(Scratch& sc, Checkpoint3<R>& k, int n) {
  int rc;
  if ((rc = sc.get(&k.id, n)) || (rc = sc.get(&k.xyz, n)) || (rc = sc.get(&k.v, n)) ||
      (rc = sc.get(&k.a, n)) || (rc = sc.get(&k.rho, n)) || (rc = sc.get(&k.u, n)) ||
      (rc = sc.get(&k.p, n)) || (rc = sc.get(&k.du, n)) || (rc = sc.get(&k.drho, n)))
    return rc;
  return ARM_OK;
}

template <typename R>
int copy_checkpoint3(arm_ctx* c, const St3<R>& S, const Checkpoint3<R>& k, bool save) {
  using R4 = typename V4<R>::type;
  const size_t n = (size_t)c->n;
  auto cp = [&](void* st, void* ck, size_t bytes) {
    return save ? hipMemcpyAsync(ck, st, bytes, hipMemcpyDeviceToDevice, c->stream)
                : hipMemcpyAsync(st, ck, bytes, hipMemcpyDeviceToDevice, c->stream);
  };
  HIPCHK(c, cp(S.id, k.id, n * sizeof(int32_t)));
  HIPCHK(c, cp(S.xyz, k.xyz, n * sizeof(double4)));
  HIPCHK(c, cp(S.v, k.v, n * sizeof(R4)));
  HIPCHK(c, cp(S.a, k.a, n * sizeof(R4)));
  HIPCHK(c, cp(S.rho, k.rho, n * sizeof(R)));
  HIPCHK(c, cp(S.u, k.u, n * sizeof(double)));
  HIPCHK(c, cp(S.p, k.p, n * sizeof(R)));
  HIPCHK(c, cp(S.du, k.du, n * sizeof(R)));
  HIPCHK(c, cp(S.drho, k.drho, n * sizeof(R)));
  return ARM_OK;
}

template <typename R> int alloc_tape3(Scratch& sc, Tape3<R>& t, int n, int ncell) {
  int rc;
  t.pf = nullptr;
  if (sizeof(R) == 4 && (rc = sc.get(&t.pf, n))) return rc;
  if ((rc = sc.get(&t.id, n)) || (rc = sc.get(&t.xyz, n)) || (rc = sc.get(&t.v, n)) ||
      (rc = sc.get(&t.rho, n)) || (rc = sc.get(&t.u, n)) || (rc = sc.get(&t.psi, n)) ||
      (rc = sc.get(&t.divv, n)) || (rc = sc.get(&t.p, n)) || (rc = sc.get(&t.drho, n)) ||
      (rc = sc.get(&t.du, n)) || (rc = sc.get(&t.a, n)) || (rc = sc.get(&t.av, n)) ||
      (rc = sc.get(&t.rho_min, n)) || (rc = sc.get(&t.u_min, n)) || (rc = sc.get(&t.v_min, n)) ||
      (rc = sc.get(&t.flag, n)) || (rc = sc.get(&t.skey, n)) || (rc = sc.get(&t.cstart, ncell + 1)) ||
      (rc = sc.get(&t.nbm, (size_t)M3_TOTAL * n)))
    return rc;
  return ARM_OK;
}

template <typename R>
void adj_force3(arm_ctx* c, const Tape3<R>& T, const AdjScratch3<R>& W, double* apart) {
  const int f = c->prm.flags;
  const int code = ((f & ARM_ART_VISC) ? 4 : 0) | ((f & ARM_ART_HEAT) ? 2 : 0) | ((f & ARM_AV_VEL) ? 1 : 0);
  const dim3 g(grid_of(c->n)), b(BLOCK);
  switch (code) {
    case 0: hipLaunchKernelGGL((k3_adj_force<R, false, false, false>), g, b, 0, c->stream, T, c->K, W, apart); break;
    case 1: hipLaunchKernelGGL((k3_adj_force<R, false, false, true>), g, b, 0, c->stream, T, c->K, W, apart); break;
    case 2: hipLaunchKernelGGL((k3_adj_force<R, false, true, false>), g, b, 0, c->stream, T, c->K, W, apart); break;
    case 3: hipLaunchKernelGGL((k3_adj_force<R, false, true, true>), g, b, 0, c->stream, T, c->K, W, apart); break;
    case 4: hipLaunchKernelGGL((k3_adj_force<R, true, false, false>), g, b, 0, c->stream, T, c->K, W, apart); break;
    case 5: hipLaunchKernelGGL((k3_adj_force<R, true, false, true>), g, b, 0, c->stream, T, c->K, W, apart); break;
    case 6: hipLaunchKernelGGL((k3_adj_force<R, true, true, false>), g, b, 0, c->stream, T, c->K, W, apart); break;
    default: hipLaunchKernelGGL((k3_adj_force<R, true, true, true>), g, b, 0, c->stream, T, c->K, W, apart); break;
  }
}

// The 3D counterpart of rollout_forward: checkpoints every ckpt steps and
// the objective, on the 3D state.
template <typename Rf>
int rollout_forward3(arm_ctx* c, int nsteps, int ckpt, const int32_t* obj_ids, int nobj,
                     double* J, double* ms) {
  const int n = c->n, ncell = c->K.ncell;
  const int nb = grid_of(n);
  drop_adjoint(c);
  AdjRun* R = new AdjRun(c);
  c->adj = R;
  Scratch& sc = R->sc;
  int rc;
  R->nsteps = nsteps;
  R->ckpt = ckpt;
  R->nck = (nsteps + ckpt - 1) / ckpt;
  R->kind = c->objkind;
  if ((rc = sc.get(&R->obj, n)) || (rc = sc.get(&R->jpart, nb)) || (rc = sc.get(&R->apart, nb)))
    return rc;
  {
    std::vector<uint8_t> h(n, 0);
    for (int k = 0; k < nobj; ++k) {
      if (obj_ids[k] < 0 || obj_ids[k] >= n)
        return fail(c, ARM_ERR_ARG, "objective id %d out of range", obj_ids[k]);
      h[obj_ids[k]] = 1;
    }
    HIPCHK(c, hipMemcpyAsync(R->obj, h.data(), n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  HIPCHK(c, hipMemsetAsync(R->jpart, 0, sizeof(double) * nb, c->stream));
  AdjArrays3<Rf>& Ar = R->arrays3<Rf>();
  Ar.ck.resize(R->nck + 1);
  for (auto& k : Ar.ck)
    if ((rc = alloc_checkpoint3(sc, k, n))) return rc;
  Ar.tape.resize(ckpt);
  for (auto& t : Ar.tape)
    if ((rc = alloc_tape3(sc, t, n, ncell))) return rc;
  AdjState3& L = R->L3;
  double** lf[] = {&L.x,   &L.y,   &L.z,    &L.vx,  &L.vy,  &L.vz,  &L.rho,   &L.u,
                   &L.bx,  &L.by,  &L.bz,   &L.bvx, &L.bvy, &L.bvz, &L.brho,  &L.bu,
                   &L.cax, &L.cay, &L.caz,  &L.cdrho, &L.cdu};
  for (double** q : lf)
    if ((rc = sc.get(q, n))) return rc;
  AdjScratch3<Rf>& W = Ar.W;
  Rf** wf[] = {&W.lax, &W.lay, &W.laz, &W.ldrho, &W.ldu, &W.lavx, &W.lavy, &W.lavz,
               &W.lp,  &W.lQ,  &W.lA,  &W.lpsi,  &W.lS,  &W.lrhoh, &W.lu, &W.lvx,
               &W.lvy, &W.lvz, &W.lx,  &W.ly,    &W.lz,  &W.lus};
  for (Rf** q : wf)
    if ((rc = sc.get(q, n))) return rc;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  zero_counts(c);
  hipLaunchKernelGGL(k3_keys<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, st3A<Rf>(c), c->K,
                     key_out(c), c->d_err);
  if (!(c->prm.flags & ARM_AV_VEL) && (rc = zero_av3<Rf>(c))) return rc;
  StepHooks fwd;
  fwd.objmask = R->obj;
  fwd.jpart = R->jpart;
  fwd.kind = R->kind;
  for (int its = 1; its <= nsteps; ++its) {
    if ((its - 1) % ckpt == 0)
      if ((rc = copy_checkpoint3(c, st3A<Rf>(c), Ar.ck[(its - 1) / ckpt], true))) return rc;
    if ((rc = fused_step3<Rf>(c, its, fwd))) return rc;
  }
  if ((rc = copy_checkpoint3(c, st3A<Rf>(c), Ar.ck[R->nck], true))) return rc;
  HIPCHK(c, hipEventRecord(e1, c->stream));
  std::vector<double> jp(nb);
  HIPCHK(c, hipMemcpyAsync(jp.data(), R->jpart, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float f = 0.f;
  (void)hipEventElapsedTime(&f, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  double Js = 0;
  for (int k = 0; k < nb; ++k) Js += jp[k];
  if (J) *J = Js;
  if (ms) *ms = f;
  c->sorted = false;
  c->masks_valid = false;
  c->have_rates = true;
  return check_domain(c);
}

template <typename Rf>
int rollout_backward3(arm_ctx* c, double seed, double* dJ_dalpha, double* dJ_dv0, double* ms) {
  AdjRun* R = c->adj;
  const int n = c->n, nb = grid_of(n);
  const int nsteps = R->nsteps, ckpt = R->ckpt, nck = R->nck;
  AdjState3& L = R->L3;
  AdjArrays3<Rf>& Ar = R->arrays3<Rf>();
  AdjScratch3<Rf>& W = Ar.W;
  int rc;
  double** lf[] = {&L.x,  &L.y,   &L.z,   &L.vx,  &L.vy,    &L.vz,  &L.rho,
                   &L.u,  &L.cax, &L.cay, &L.caz, &L.cdrho, &L.cdu};
  for (double** q : lf) HIPCHK(c, hipMemsetAsync(*q, 0, sizeof(double) * n, c->stream));
  HIPCHK(c, hipMemsetAsync(R->apart, 0, sizeof(double) * nb, c->stream));
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  for (int seg = nck - 1; seg >= 0; --seg) {
    const int s0 = 1 + seg * ckpt, s1 = std::min(s0 + ckpt, nsteps + 1);
    if ((rc = copy_checkpoint3(c, st3A<Rf>(c), Ar.ck[seg], false))) return rc;
    zero_counts(c);
    hipLaunchKernelGGL(k3_keys<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, st3A<Rf>(c), c->K,
                       key_out(c), c->d_err);
    for (int its = s0; its < s1; ++its) {
      StepHooks rec;
      rec.tape = &Ar.tape[its - s0];
      if ((rc = fused_step3<Rf>(c, its, rec))) return rc;
    }
    for (int its = s1 - 1; its >= s0; --its) {
      const Tape3<Rf>& T = Ar.tape[its - s0];
      hipLaunchKernelGGL(k3_adj_corrector<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, its,
                         L, W, R->obj, R->kind, seed);
      adj_force3<Rf>(c, T, W, R->apart);
      hipLaunchKernelGGL(k3_adj_eos<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, W);
      hipLaunchKernelGGL(k3_adj_density<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, W);
      hipLaunchKernelGGL(k3_adj_predictor<Rf>, dim3(nb), dim3(BLOCK), 0, c->stream, T, c->K, its,
                         L, W);
      HIPCHK(c, hipGetLastError());
      std::swap(L.x, L.bx);
      std::swap(L.y, L.by);
      std::swap(L.z, L.bz);
      std::swap(L.vx, L.bvx);
      std::swap(L.vy, L.bvy);
      std::swap(L.vz, L.bvz);
      std::swap(L.rho, L.brho);
      std::swap(L.u, L.bu);
    }
  }
  HIPCHK(c, hipEventRecord(e1, c->stream));
  if ((rc = copy_checkpoint3(c, st3A<Rf>(c), Ar.ck[nck], false))) return rc;
  std::vector<double> ap(nb), gx(n), gy(n), gz(n);
  HIPCHK(c, hipMemcpyAsync(ap.data(), R->apart, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(gx.data(), L.vx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(gy.data(), L.vy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(gz.data(), L.vz, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float f = 0.f;
  (void)hipEventElapsedTime(&f, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  double As = 0;
  for (int k = 0; k < nb; ++k) As += ap[k];
  c->grads[ARM_GRAD_ALPHA] = As;
  c->grads[ARM_GRAD_Q0] = 0.0;
  c->grads[ARM_GRAD_VISC] = 0.0;
  if (dJ_dalpha) *dJ_dalpha = As;
  if (dJ_dv0)
    for (int i = 0; i < n; ++i) {
      dJ_dv0[3 * i] = gx[i];
      dJ_dv0[3 * i + 1] = gy[i];
      dJ_dv0[3 * i + 2] = gz[i];
    }
  if (ms) *ms = f;
  c->sorted = false;
  c->masks_valid = false;
  return check_domain(c);
}

#define DISPATCH(c, call) (is32(c) ? call<float> : call<double>)

void set_kparams(KParams& K, const arm_params& P) {
  K.flags = P.flags;
  K.all_rate = (P.flags & ARM_DBC) ? 1 : 0;
  K.pa_sph = P.pa_sph;
  K.density_method = P.density_method;
  K.nx = P.nx;
  K.ny = P.ny;
  K.ncell = P.nx * P.ny;
  K.h = P.h;
  K.invh = 1.0 / P.h;
  K.mass = P.mass;
  K.rho0 = P.rho0;
  K.c0 = P.c0;
  K.pmin = P.pmin;
  K.B = P.rho0 * P.c0 * P.c0 / 7.0;
  K.av_alpha = P.av_alpha;
  K.av_beta = P.av_beta;
  K.av_e2 = (P.h * P.av_eta) * (P.h * P.av_eta);
  K.ah_g1 = P.ah_g1;
  K.ah_g2 = P.ah_g2;
  K.ah_e2 = (P.h * P.ah_eta) * (P.h * P.ah_eta);
  K.xsph_eps = P.xsph_eps;
  K.gx = P.gx;
  K.gy = P.gy;
  K.dt = P.dt;
  K.ox = P.ox;
  K.oy = P.oy;
  K.cell = P.cell;
  K.cut2 = (2.0 * P.h) * (2.0 * P.h);
  K.wnorm = 7.0 / (4.0 * M_PI * P.h * P.h);
  K.C1 = 0.625 * K.wnorm / (P.h * P.h);
  K.psi0 = P.ah_g1 * P.h * P.c0;
  K.g2h2 = P.ah_g2 * P.h * P.h;
  K.eps_m = P.xsph_eps * P.mass;
  K.same_e2 = (K.av_e2 == K.ah_e2) ? 1 : 0;
  K.f_cm = -8.0 * K.C1 * P.mass;
  K.f_ava = -0.5 * P.av_alpha * P.c0;
  K.f_avb = 0.5 * P.av_beta;
  K.f_xs = 0.5 * K.wnorm * K.eps_m;
  K.cellf = (float)P.cell;
  K.cut2f = (float)K.cut2;
  {
    // x_a = fl(cx cell + off) is within u |x_a| <= u nx cell of the exact
    // value (u = 2^-24); a difference of two, rounded, and the y and z terms
    // (oy_i + uy) - oy_j rounded in another order, are within
    // E = u (2 nx cell + 4 cell + 16 h) of the exact path's per axis, so
    // r < 2h implies r_a < 2h + 2E
    const double u = 1.0 / 16777216.0;
    const double nmax = (double)std::max(P.nx, std::max(P.ny, P.dim == 3 ? P.nz : 1));
    const double E = u * (2.0 * nmax * P.cell + 4.0 * P.cell + 16.0 * P.h);
    const double b = 2.0 * P.h + 2.0 * E;
    K.pre2f = std::nextafter((float)(b * b * (1.0 + 1e-5)), INFINITY);
  }
  K.eos = P.eos;
  K.lin_c2 = P.c0 * P.c0;
  K.lin_p0 = P.eos_p0;
  K.lin_p0r = P.eos_p0 + K.lin_c2 * P.rho0;
  K.eta = P.visc_eta;
  K.hs_q0 = P.hs_q0;
  K.hs_x = P.hs_x;
  K.hs_y = P.hs_y;
  K.hs_k = P.hs_sigma > 0 ? 1.0 / (2.0 * P.hs_sigma * P.hs_sigma) : 0.0;
  K.sle = P.sle;
  K.av_eta = P.av_eta;
  K.ah_eta = P.ah_eta;
  K.periodic = (P.flags & ARM_PERIODIC_X) ? 1 : 0;
  K.cond = (P.flags & ARM_CONDUCTION)
               ? 4.0 * P.cond_k * P.cond_k / (P.cond_k + P.cond_k) / P.cp
               : 0.0;
  K.f_cond2 = 2.0 * K.cond;
  K.contact_h = (P.flags & ARM_CONTACT) ? P.h / P.contact_hdr : 0.0;
  K.dim = P.dim == 3 ? 3 : 2;
  K.nz = K.dim == 3 ? P.nz : 1;
  K.oz = P.oz;
  K.gz = P.gz;
  if (K.dim == 3) {
    K.ncell = P.nx * P.ny * P.nz;
    K.wnorm = 21.0 / (16.0 * M_PI * P.h * P.h * P.h);
    K.C1 = 0.625 * K.wnorm / (P.h * P.h);
    K.f_cm = -8.0 * K.C1 * P.mass;
    K.f_xs = 0.5 * K.wnorm * K.eps_m;
  }
  set_kf(K);
}

int need_sorted(arm_ctx* c) {
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (!c->sorted) return fail(c, ARM_ERR_STATE, "arm_link_list has not run on this state");
  return ARM_OK;
}

}  // namespace

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

int arm_create(const arm_params* prm, int capacity, arm_ctx** out) {
  if (!prm || !out || capacity <= 0) return ARM_ERR_ARG;
  *out = nullptr;
  arm_ctx* c = new (std::nothrow) arm_ctx();
  if (!c) return ARM_ERR_ARG;
  c->prm = *prm;
  const arm_params& P = *prm;
  int rc = ARM_OK;
  if (P.precision != ARM_FP64 && P.precision != ARM_FP32)
    rc = fail(c, ARM_ERR_ARG, "precision must be ARM_FP64 or ARM_FP32");
  else if (!(P.h > 0) || !(P.mass > 0) || !(P.rho0 > 0) || !(P.dt > 0))
    rc = fail(c, ARM_ERR_ARG, "h, mass, rho0 and dt must be positive");
  else if (!(P.cell >= 2.0 * P.h))
    rc = fail(c, ARM_ERR_ARG, "cell (%g) must be at least the kernel support 2h (%g)", P.cell, 2 * P.h);
  else if (P.dim != 0 && P.dim != 2 && P.dim != 3)
    rc = fail(c, ARM_ERR_ARG, "dim must be 2 or 3");
  else if (P.nx < 1 || P.ny < 1 || (long long)P.nx * P.ny >= (1LL << 30))
    rc = fail(c, ARM_ERR_ARG, "bad grid %d x %d", P.nx, P.ny);
  else if (P.dim == 3 && (P.nz < 1 || (long long)P.nx * P.ny * P.nz >= (1LL << 30)))
    rc = fail(c, ARM_ERR_ARG, "bad grid %d x %d x %d", P.nx, P.ny, P.nz);
  else if (P.dim != 3 && P.precision == ARM_FP32 && (P.nx > 65535 || P.ny > 65535))
    rc = fail(c, ARM_ERR_ARG, "FP32 form packs cell coordinates in 16 bits");
  else if (P.dim == 3 && (P.flags & (ARM_VISC | ARM_HEAT_SOURCE | ARM_CONDUCTION | ARM_PERIODIC_X)))
    rc = fail(c, ARM_ERR_UNSUPPORTED,
              "ARM_VISC, ARM_HEAT_SOURCE, ARM_CONDUCTION and ARM_PERIODIC_X are 2D only");
  else if (P.sle < 0 || P.sle > 2)
    rc = fail(c, ARM_ERR_ARG, "sle must be 0, 1 or 2");
  else if (P.dim == 3 && P.sle)
    rc = fail(c, ARM_ERR_UNSUPPORTED, "the smoothing-length evolution is 2D only");
  else if ((P.flags & ARM_PERIODIC_X) && P.nx < 3)
    rc = fail(c, ARM_ERR_ARG, "a periodic grid needs nx >= 3 cells");
  else if ((P.flags & ARM_CONDUCTION) && !(P.cond_k > 0 && P.cp > 0))
    rc = fail(c, ARM_ERR_ARG, "conduction needs cond_k > 0 and cp > 0");
  else if (P.pa_sph != 0 && P.pa_sph != 1)
    rc = fail(c, ARM_ERR_UNSUPPORTED, "pa_sph must be 0 or 1");
  else if (P.density_method != 0 && P.density_method != 1)
    rc = fail(c, ARM_ERR_UNSUPPORTED, "density_method must be 0 (continuity) or 1 (summation)");
  else if (P.eos != ARM_EOS_TAIT && P.eos != ARM_EOS_LINEAR)
    rc = fail(c, ARM_ERR_UNSUPPORTED, "eos must be ARM_EOS_TAIT or ARM_EOS_LINEAR");
  else if ((P.flags & ARM_VISC) && !(P.visc_eta >= 0))
    rc = fail(c, ARM_ERR_ARG, "visc_eta must be non-negative");
  else if ((P.flags & ARM_HEAT_SOURCE) && !(P.hs_sigma > 0))
    rc = fail(c, ARM_ERR_ARG, "hs_sigma must be positive");
  else if ((P.flags & ARM_GHOSTS) &&
           (P.dim == 3 || (P.flags & (ARM_DBC | ARM_PERIODIC_X)) || P.sle))
    rc = fail(c, ARM_ERR_UNSUPPORTED,
              "ARM_GHOSTS is 2D, without ARM_DBC, ARM_PERIODIC_X or sle");
  else if ((P.flags & ARM_CONTACT) &&
           (P.dim == 3 || (P.flags & (ARM_DBC | ARM_GHOSTS | ARM_PERIODIC_X)) || P.sle))
    rc = fail(c, ARM_ERR_UNSUPPORTED,
              "ARM_CONTACT is 2D, without ARM_DBC, ARM_GHOSTS, ARM_PERIODIC_X or sle");
  else if ((P.flags & ARM_CONTACT) && !(P.contact_hdr > 0))
    rc = fail(c, ARM_ERR_ARG, "ARM_CONTACT needs contact_hdr > 0");
  if (rc) {
    *out = c;  // so that arm_last_error can report
    return rc;
  }
  set_kparams(c->K, P);

  c->cap = capacity;
  const int ncell = c->K.ncell;
  int nbits = 1;
  while (nbits < 31 && (1LL << nbits) < (long long)ncell) ++nbits;
  c->sort_passes = (nbits + SORT_MAX_BITS - 1) / SORT_MAX_BITS;
  c->sort_bits = (nbits + c->sort_passes - 1) / c->sort_passes;
  c->ntiles = (capacity + SORT_TILE - 1) / SORT_TILE;

  hipError_t e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) rc = fail(c, ARM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  if (is3d(c)) {
    if (!rc) rc = is32(c) ? alloc_state3(c, c->A3_32, capacity) : alloc_state3(c, c->A3_64, capacity);
    if (!rc) rc = is32(c) ? alloc_state3(c, c->B3_32, capacity) : alloc_state3(c, c->B3_64, capacity);
  } else {
    if (!rc) rc = is32(c) ? alloc_state(c, c->A32, capacity) : alloc_state(c, c->A64, capacity);
    if (!rc) rc = is32(c) ? alloc_state(c, c->B32, capacity) : alloc_state(c, c->B64, capacity);
  }
  if (!rc) rc = dalloc(c, &c->key_by_id, capacity);
  if (!rc) rc = dalloc(c, &c->slot_by_id, capacity);
  for (int k = 0; k < 2 && !rc; ++k) {
    rc = dalloc(c, &c->sk[k], capacity);
    if (!rc) rc = dalloc(c, &c->sv[k], capacity);
  }
  if (!rc) rc = dalloc(c, &c->hist, (size_t)(1 << c->sort_bits) * c->ntiles);
  if (!rc) rc = dalloc(c, &c->tot, (size_t)(1 << c->sort_bits));
  c->ghosts = (P.flags & ARM_GHOSTS) != 0;
  // the counting sort runs over ncell cells, plus one with ghosts; its tile
  // sums also serve the ghost numbering (a scan over the real ids)
  c->cs_tiles = (ncell + 1 + CS_TILE - 1) / CS_TILE;
  c->counting = c->cs_tiles <= CS_MAX_TILES;
  if (!rc && c->ghosts && !c->counting)
    rc = fail(c, ARM_ERR_UNSUPPORTED, "ARM_GHOSTS needs a grid of at most %d cells",
              CS_TILE * CS_MAX_TILES - 1);
  if (!rc && c->counting) {
    const int tiles = std::max(c->cs_tiles, (capacity + CS_TILE - 1) / CS_TILE);
    rc = dalloc(c, &c->cnt, (size_t)ncell + 1);
    if (!rc) rc = dalloc(c, &c->tsum, (size_t)tiles);
    if (!rc) rc = dalloc(c, &c->tstat, (size_t)c->cs_tiles);
    if (!rc) {
      hipError_t ez = hipMemset(c->tstat, 0, sizeof(uint64_t) * (size_t)c->cs_tiles);
      if (ez != hipSuccess) rc = fail(c, ARM_ERR_HIP, "hipMemset: %s", hipGetErrorString(ez));
    }
    if (!rc) rc = dalloc(c, &c->ent, (size_t)capacity);
    if (!rc) {
      hipError_t ez = hipMemset(c->cnt, 0, sizeof(int) * ((size_t)ncell + 1));
      if (ez != hipSuccess) rc = fail(c, ARM_ERR_HIP, "hipMemset: %s", hipGetErrorString(ez));
    }
  }
  if (!rc) rc = dalloc(c, &c->cstart, (size_t)ncell + 2);
  if (!rc && c->ghosts) {
    Ghosts& G = c->G;
    rc = dalloc(c, &G.xprev, capacity);
    if (!rc) rc = dalloc(c, &G.gsel, capacity);
    if (!rc) rc = dalloc(c, &G.gflag, capacity);
    if (!rc) rc = dalloc(c, &G.grank, (size_t)capacity + 1);
    if (!rc) rc = dalloc(c, &G.gpar, capacity);
    if (!rc) rc = dalloc(c, &G.gwall, capacity);
    if (!rc) rc = dalloc(c, &G.gpid, capacity);
    if (!rc) rc = dalloc(c, &G.sid, capacity);
    if (!rc) rc = dalloc(c, &G.vnear, (size_t)ncell);
    if (!rc) rc = dalloc(c, &G.vstart, (size_t)ncell + 1);
    if (!rc) rc = dalloc(c, &G.vxy, capacity);
    if (!rc) rc = dalloc(c, &G.vnrm, capacity);
    if (!rc) rc = dalloc(c, &G.vid, capacity);
    if (!rc) {
      void* hp = nullptr;
      e = hipHostMalloc(&hp, 4 * sizeof(int), hipHostMallocDefault);
      if (e != hipSuccess) rc = fail(c, ARM_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
      c->h_pin = static_cast<int*>(hp);
    }
  }
  if (!rc && (P.flags & ARM_CONTACT)) {
    double2* q = nullptr;
    rc = dalloc(c, &q, capacity);
    c->d_cd = q;
  }
  if (!rc) rc = dalloc(c, &c->d_err, 4);
  if (!rc) c->d_ng = c->d_err + 2;
  if (!rc) rc = dalloc(c, &c->d_part, (size_t)grid_of(capacity) * 6);
  if (!rc) rc = dalloc(c, &c->d_cnt, capacity);
  if (!rc) rc = dalloc(c, &c->nbm, capacity);
  if (!rc && !is3d(c)) {
    rc = dalloc(c, &c->pieces, (size_t)capacity / PIECE + c->K.ny + 2);
    if (!rc) rc = dalloc(c, &c->npieces, 1);
  }
  if (!rc && is3d(c)) rc = dalloc(c, &c->nbm3, (size_t)M3_TOTAL * capacity);
  if (!rc && is3d(c)) {
    const size_t r4 = is32(c) ? sizeof(float4) : sizeof(double4);
    double4* q = nullptr;
    rc = dalloc(c, &q, (3 * r4 * capacity + sizeof(double4) - 1) / sizeof(double4));
    c->pack3_vq = q;
  }
  if (!rc) {
    e = hipMemset(c->d_err, 0, 4 * sizeof(int));
    if (e != hipSuccess) rc = fail(c, ARM_ERR_HIP, "hipMemset: %s", hipGetErrorString(e));
  }
  if (!rc) {
    e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) rc = fail(c, ARM_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  *out = c;
  return rc;
}

void arm_destroy(arm_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  resolve_times(c);
  drop_adjoint(c);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->d_dep) (void)hipFree(c->d_dep);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* arm_last_error(const arm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int arm_domain_violations(const arm_ctx* c) { return c ? c->domain_bad : 0; }

int arm_upload(arm_ctx* c, int np, int nv, const double* x, const double* v, const double* rho,
               const double* u) {
  if (!c || !x || np < 0 || nv < 0) return ARM_ERR_ARG;
  if (np + nv > c->cap) return fail(c, ARM_ERR_CAPACITY, "%d particles exceed capacity %d", np + nv, c->cap);
  if (np + nv < 1) return fail(c, ARM_ERR_ARG, "no particles");
  c->n = np + nv;
  c->np = np;
  c->nb = c->n;
  c->ng = c->ng_new = 0;
  c->ghost_pending = false;
  c->have_ranges = false;
  c->K.n = c->n;
  c->K.np = np;
  c->K.nb = c->n;
  c->sorted = false;
  c->masks_valid = false;
  c->have_rates = false;
  c->domain_bad = 0;
  int rc = is3d(c) ? DISPATCH(c, upload3_t)(c, x, v, rho, u)
                    : DISPATCH(c, upload_t)(c, x, v, rho, nullptr, u);
  if (!rc && c->ghosts) {  // no search yet: the previous positions are these
    const double2* xy = is32(c) ? c->A32.xy : c->A64.xy;
    hipError_t e = hipMemcpyAsync(c->G.xprev, xy, sizeof(double2) * (size_t)c->n,
                                  hipMemcpyDeviceToDevice, c->stream);
    // the virtual particles never move: their (cell, id) list, once
    {
      const int ncell = c->K.ncell;
      std::vector<std::pair<long long, int>> kv;
      kv.reserve(nv);
      for (int j = np; j < np + nv; ++j) {
        const double tx = (x[2 * j] - c->K.ox) / c->K.cell;
        const double ty = (x[2 * j + 1] - c->K.oy) / c->K.cell;
        long long key = ncell;  // outside the grid: never a candidate
        if (tx >= 0.0 && ty >= 0.0 && tx < (double)c->K.nx && ty < (double)c->K.ny)
          key = (long long)(int)tx + (long long)(int)ty * c->K.nx;
        kv.emplace_back(key, j);
      }
      std::sort(kv.begin(), kv.end());
      std::vector<int> vstart((size_t)ncell + 1, 0), vid(nv);
      std::vector<double2> vxy(nv), vnrm(nv);
      size_t q = 0;
      for (int cc = 0; cc <= ncell; ++cc) {
        while (q < kv.size() && kv[q].first < cc) ++q;
        vstart[cc] = (int)q;
      }
      for (int k = 0; k < nv; ++k) {
        const int j = kv[k].second;
        vid[k] = j;
        vxy[k] = make_double2(x[2 * j], x[2 * j + 1]);
        vnrm[k] = make_double2(v ? v[2 * j] : 0.0, v ? v[2 * j + 1] : 0.0);
      }
      if (e == hipSuccess)
        e = hipMemcpyAsync(c->G.vstart, vstart.data(), sizeof(int) * vstart.size(),
                           hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess && nv > 0)
        e = hipMemcpyAsync(c->G.vid, vid.data(), sizeof(int) * nv, hipMemcpyHostToDevice,
                           c->stream);
      if (e == hipSuccess && nv > 0)
        e = hipMemcpyAsync(c->G.vxy, vxy.data(), sizeof(double2) * nv, hipMemcpyHostToDevice,
                           c->stream);
      if (e == hipSuccess && nv > 0)
        e = hipMemcpyAsync(c->G.vnrm, vnrm.data(), sizeof(double2) * nv,
                           hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    // the cells near the (fixed) virtual particles
    if (e == hipSuccess) e = hipMemsetAsync(c->G.vnear, 0, (size_t)c->K.ncell, c->stream);
    if (e == hipSuccess) {
      if (is32(c))
        hipLaunchKernelGGL(k_ghost_mark<float>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream,
                           c->A32, c->K, c->G, c->n);
      else
        hipLaunchKernelGGL(k_ghost_mark<double>, dim3(grid_of(c->n)), dim3(BLOCK), 0,
                           c->stream, c->A64, c->K, c->G, c->n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(c, ARM_ERR_HIP, "ghost set-up: %s", hipGetErrorString(e));
  }
  c->loaded = (rc == ARM_OK);
  c->its_next = 1;
  return rc;
}

int arm_download(arm_ctx* c, double* x, double* v, double* rho, double* p, double* u) {
  if (!c) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (is3d(c)) return DISPATCH(c, download3_t)(c, x, v, rho, p, u);
  return DISPATCH(c, download_t)(c, x, v, rho, p, u);
}

int arm_download_rates(arm_ctx* c, double* drho, double* dvdt, double* dudt, double* av,
                       double* divv) {
  if (!c) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (is3d(c)) return DISPATCH(c, download_rates3_t)(c, drho, dvdt, dudt, av, divv);
  return DISPATCH(c, download_rates_t)(c, drho, dvdt, dudt, av, divv);
}

int arm_upload_h(arm_ctx* c, const double* h) {
  if (!c || !h) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (!c->prm.sle || is3d(c)) return fail(c, ARM_ERR_STATE, "per-particle h needs sle = 1 or 2 (2D)");
  const int n = c->n;
  std::vector<int32_t> id(n);
  const int32_t* did = is32(c) ? c->A32.id : c->A64.id;
  double* dh = is32(c) ? c->A32.hs : c->A64.hs;
  HIPCHK(c, hipMemcpyAsync(id.data(), did, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> hh(n);
  for (int s = 0; s < n; ++s) {
    const double v = h[id[s]];
    if (!(v > 0) || !(2.0 * v <= c->prm.cell))
      return fail(c, ARM_ERR_ARG, "h[%d] = %g outside (0, cell/2]", id[s], v);
    hh[s] = v;
  }
  HIPCHK(c, hipMemcpyAsync(dh, hh.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ARM_OK;
}

int arm_download_h(arm_ctx* c, double* h) {
  if (!c || !h) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (is3d(c)) return fail(c, ARM_ERR_UNSUPPORTED, "the smoothing-length evolution is 2D only");
  const int n = c->n;
  std::vector<int32_t> id(n);
  std::vector<double> hh(n);
  const int32_t* did = is32(c) ? c->A32.id : c->A64.id;
  const double* dh = is32(c) ? c->A32.hs : c->A64.hs;
  HIPCHK(c, hipMemcpyAsync(id.data(), did, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hh.data(), dh, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < n; ++s) h[id[s]] = hh[s];
  return ARM_OK;
}

int arm_link_list(arm_ctx* c, int* niac) {
  if (!c) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  int rc = is3d(c) ? DISPATCH(c, link_list3_t)(c) : DISPATCH(c, link_list_t)(c);
  if (rc) return rc;
  if (niac)
    rc = is3d(c) ? DISPATCH(c, pairs3_t)(c, 0, niac, nullptr, nullptr)
                 : DISPATCH(c, pairs_t)(c, 0, niac, nullptr, nullptr);
  return rc;
}

int arm_get_pairs(arm_ctx* c, int cap, int* niac, int* pair_i, int* pair_j) {
  if (!c || !pair_i || !pair_j || cap < 0) return ARM_ERR_ARG;
  int rc = need_sorted(c);
  if (rc) return rc;
  if (is3d(c)) return DISPATCH(c, pairs3_t)(c, cap, niac, pair_i, pair_j);
  return DISPATCH(c, pairs_t)(c, cap, niac, pair_i, pair_j);
}

int arm_get_cells(arm_ctx* c, int32_t* key_sorted, int32_t* id_sorted, int32_t* cell_start) {
  if (!c) return ARM_ERR_ARG;
  int rc = need_sorted(c);
  if (rc) return rc;
  const int32_t* dids = is3d(c) ? (is32(c) ? c->A3_32.id : c->A3_64.id)
                                : (is32(c) ? c->A32.id : c->A64.id);
  if (key_sorted)
    HIPCHK(c, hipMemcpyAsync(key_sorted, c->skey, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
  if (id_sorted)
    HIPCHK(c, hipMemcpyAsync(id_sorted, dids, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
  if (cell_start)
    HIPCHK(c, hipMemcpyAsync(cell_start, c->cstart, sizeof(int32_t) * ((size_t)c->K.ncell + 1),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ARM_OK;
}

int arm_density(arm_ctx* c, int method) {
  if (!c) return ARM_ERR_ARG;
  if (method != 0 && method != 1) return fail(c, ARM_ERR_UNSUPPORTED, "density method %d", method);
  int rc = need_sorted(c);
  if (rc) return rc;
  rc = is3d(c) ? DISPATCH(c, density3_t)(c, method, false)
               : DISPATCH(c, density_t)(c, method, false);
  if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

int arm_eos(arm_ctx* c) {
  if (!c) return ARM_ERR_ARG;
  int rc = need_sorted(c);
  if (rc) return rc;
  rc = is3d(c) ? DISPATCH(c, eos3_t)(c) : DISPATCH(c, eos_t)(c);
  if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

int arm_internal_force(arm_ctx* c, int flags) {
  if (!c) return ARM_ERR_ARG;
  int rc = need_sorted(c);
  if (rc) return rc;
  rc = is3d(c) ? DISPATCH(c, force3_t)(c, true, flags, false)
               : DISPATCH(c, force_t)(c, true, flags, false);
  if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

int arm_av_vel(arm_ctx* c) {
  if (!c) return ARM_ERR_ARG;
  int rc = need_sorted(c);
  if (rc) return rc;
  rc = is3d(c) ? DISPATCH(c, force3_t)(c, false, 0, true)
               : DISPATCH(c, force_t)(c, false, 0, true);
  if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

int arm_single_step(arm_ctx* c) {
  if (!c) return ARM_ERR_ARG;
  int rc = need_sorted(c);
  if (rc) return rc;
  rc = is3d(c) ? DISPATCH(c, single_step3_t)(c) : DISPATCH(c, single_step_t)(c);
  if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

int arm_predictor(arm_ctx* c) {
  if (!c) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  int rc = is3d(c) ? DISPATCH(c, predictor3_t)(c) : DISPATCH(c, predictor_t)(c);
  if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

int arm_corrector(arm_ctx* c, int its) {
  if (!c) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (!c->have_rates) return fail(c, ARM_ERR_STATE, "no rates: run arm_single_step first");
  int rc = is3d(c) ? DISPATCH(c, corrector3_t)(c, its, false)
                    : DISPATCH(c, corrector_t)(c, its, false);
  if (!rc) HIPCHK(c, hipStreamSynchronize(c->stream));
  // positions moved: the cell order and masks must be rebuilt before the next gather
  c->masks_valid = false;
  if (!rc && is32(c) && is3d(c)) {
    hipLaunchKernelGGL(k3_keys<float>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, c->A3_32,
                       c->K, KeyOut{}, c->d_err);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else if (!rc && is32(c)) {
    // refresh cell + offset of the FP32 form from the FP64 accumulators
    St<float>& A = c->A32;
    hipLaunchKernelGGL(k_keys<float>, dim3(grid_of(c->n)), dim3(BLOCK), 0, c->stream, A, c->K,
                       KeyOut{}, c->d_err, 0);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->sorted = false;
  return rc ? rc : check_domain(c);
}

int arm_time_integration(arm_ctx* c, int its0, int nsteps) {
  if (!c || its0 < 1 || nsteps < 0) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (nsteps == 0) return ARM_OK;
  int rc = is3d(c) ? DISPATCH(c, time_integration3_t)(c, its0, nsteps)
                    : DISPATCH(c, time_integration_t)(c, its0, nsteps);
  // the last corrector moved the particles after the last sort
  c->sorted = false;
  return rc;
}

int arm_ghost_particles(arm_ctx* c, int* ng) {
  if (!c) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (!c->ghosts) return fail(c, ARM_ERR_UNSUPPORTED, "the context was created without ARM_GHOSTS");
  if (!c->have_ranges) return fail(c, ARM_ERR_STATE, "arm_link_list has not run on this state");
  if (c->ghost_pending)
    return fail(c, ARM_ERR_STATE, "arm_link_list must follow arm_ghost_particles");
  int rc = DISPATCH(c, make_ghosts)(c, false);
  if (rc) return rc;
  c->gfirst = c->n;
  c->n = c->K.n = c->n + c->ng_new;
  c->ghost_pending = true;
  c->sorted = false;
  c->masks_valid = false;
  if (ng) *ng = c->ng_new;
  return ARM_OK;
}

int arm_ghost_count(arm_ctx* c, int* ng) {
  if (!c || !ng) return ARM_ERR_ARG;
  *ng = c->ghost_pending ? c->ng_new : c->ng;
  return ARM_OK;
}

int arm_download_ghosts(arm_ctx* c, double* x, double* v, double* rho, double* p, double* u,
                        int32_t* parent) {
  if (!c) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (!c->ghosts) return fail(c, ARM_ERR_UNSUPPORTED, "the context was created without ARM_GHOSTS");
  if (c->ghost_pending) return fail(c, ARM_ERR_STATE, "arm_link_list must run first");
  return DISPATCH(c, download_ghosts_t)(c, x, v, rho, p, u, parent);
}

int arm_conservation(arm_ctx* c, double out[5]) {
  if (!c || !out) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (is3d(c)) {
    double o6[6];
    int rc = DISPATCH(c, conservation3_t)(c, o6);
    if (rc) return rc;
    out[0] = o6[0];
    out[1] = o6[1];
    out[2] = o6[2];
    out[3] = o6[4];
    out[4] = o6[5];
    return ARM_OK;
  }
  return DISPATCH(c, conservation_t)(c, out);
}

int arm_conservation3(arm_ctx* c, double out[6]) {
  if (!c || !out) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (is3d(c)) return DISPATCH(c, conservation3_t)(c, out);
  double o5[5];
  int rc = DISPATCH(c, conservation_t)(c, o5);
  if (rc) return rc;
  out[0] = o5[0];
  out[1] = o5[1];
  out[2] = o5[2];
  out[3] = 0.0;
  out[4] = o5[3];
  out[5] = o5[4];
  return ARM_OK;
}

int arm_pass_times(arm_ctx* c, int n, double* ms, int* counts) {
  if (!c || n < 0) return ARM_ERR_ARG;
  resolve_times(c);
  for (int k = 0; k < n && k < ARM_PASS_COUNT; ++k) {
    if (ms) ms[k] = c->pass_ms[k];
    if (counts) counts[k] = c->pass_cnt[k];
  }
  return ARM_OK;
}

int arm_reset_pass_times(arm_ctx* c) {
  if (!c) return ARM_ERR_ARG;
  resolve_times(c);
  for (int k = 0; k < ARM_PASS_COUNT; ++k) {
    c->pass_ms[k] = 0;
    c->pass_cnt[k] = 0;
  }
  return ARM_OK;
}

int arm_rollout_forward(arm_ctx* c, int nsteps, int ckpt_every, const int32_t* obj_ids, int nobj,
                        double* J, double* ms) {
  if (!c || nsteps < 1 || ckpt_every < 1 || nobj < 0 || (nobj > 0 && !obj_ids)) return ARM_ERR_ARG;
  if (!c->loaded) return fail(c, ARM_ERR_STATE, "no state uploaded");
  if (c->prm.density_method != 0)
    return fail(c, ARM_ERR_UNSUPPORTED, "the adjoint differentiates continuity density only");
  if (c->prm.flags & (ARM_CONDUCTION | ARM_PERIODIC_X))
    return fail(c, ARM_ERR_UNSUPPORTED, "the adjoint has no conduction or periodic boundary");
  if (c->ghosts) return fail(c, ARM_ERR_UNSUPPORTED, "the adjoint has no mirror ghosts");
  if (c->d_dep) return fail(c, ARM_ERR_UNSUPPORTED, "the adjoint has no deposition heat source");
  if (c->prm.sle) return fail(c, ARM_ERR_UNSUPPORTED, "the adjoint keeps h fixed (sle = 0)");
  if (c->prm.flags & ARM_CONTACT)
    return fail(c, ARM_ERR_UNSUPPORTED, "the adjoint has no contact displacement");
  const int ck = std::min(ckpt_every, nsteps);
  int rc;
  if (is3d(c))
    rc = is32(c) ? rollout_forward3<float>(c, nsteps, ck, obj_ids, nobj, J, ms)
                 : rollout_forward3<double>(c, nsteps, ck, obj_ids, nobj, J, ms);
  else
    rc = is32(c) ? rollout_forward<float>(c, nsteps, ck, obj_ids, nobj, J, ms)
                 : rollout_forward<double>(c, nsteps, ck, obj_ids, nobj, J, ms);
  if (rc) drop_adjoint(c);
  return rc;
}

int arm_rollout_backward(arm_ctx* c, double seed, double* dJ_dalpha, double* dJ_dv0, double* ms) {
  if (!c) return ARM_ERR_ARG;
  if (!c->adj) return fail(c, ARM_ERR_STATE, "no arm_rollout_forward to differentiate");
  if (is3d(c))
    return is32(c) ? rollout_backward3<float>(c, seed, dJ_dalpha, dJ_dv0, ms)
                   : rollout_backward3<double>(c, seed, dJ_dalpha, dJ_dv0, ms);
  return is32(c) ? rollout_backward<float>(c, seed, dJ_dalpha, dJ_dv0, ms)
                 : rollout_backward<double>(c, seed, dJ_dalpha, dJ_dv0, ms);
}

int arm_adjoint_rollout(arm_ctx* c, int nsteps, int ckpt_every, const int32_t* obj_ids, int nobj,
                        double* J, double* dJ_dalpha, double* dJ_dv0, double* ms_fwd,
                        double* ms_bwd) {
  int rc = arm_rollout_forward(c, nsteps, ckpt_every, obj_ids, nobj, J, ms_fwd);
  if (rc) return rc;
  return arm_rollout_backward(c, 1.0, dJ_dalpha, dJ_dv0, ms_bwd);
}

int arm_update_params(arm_ctx* c, const arm_params* prm) {
  if (!c || !prm) return ARM_ERR_ARG;
  const arm_params& P = *prm;
  const arm_params& O = c->prm;
  if (P.precision != O.precision || P.nx != O.nx || P.ny != O.ny || P.ox != O.ox ||
      P.oy != O.oy || P.cell != O.cell || P.h != O.h || P.flags != O.flags ||
      P.density_method != O.density_method || P.pa_sph != O.pa_sph || P.eos != O.eos ||
      P.dim != O.dim || (P.dim == 3 && (P.nz != O.nz || P.oz != O.oz)) || P.sle != O.sle)
    return fail(c, ARM_ERR_ARG, "only the physical scalars can change after arm_create");
  if ((P.flags & ARM_VISC) && !(P.visc_eta >= 0))
    return fail(c, ARM_ERR_ARG, "visc_eta must be non-negative");
  if ((P.flags & ARM_HEAT_SOURCE) && !(P.hs_sigma > 0))
    return fail(c, ARM_ERR_ARG, "hs_sigma must be positive");
  if ((P.flags & ARM_CONDUCTION) && !(P.cond_k > 0 && P.cp > 0))
    return fail(c, ARM_ERR_ARG, "conduction needs cond_k > 0 and cp > 0");
  if (!(P.mass > 0) || !(P.rho0 > 0) || !(P.dt > 0))
    return fail(c, ARM_ERR_ARG, "mass, rho0 and dt must be positive");
  c->prm = P;
  set_kparams(c->K, P);
  drop_adjoint(c);
  return ARM_OK;
}

int arm_set_deposition(arm_ctx* c, const double* origin, const double* cell, const int* nbin,
                       const double* data, double t_end) {
  return arm_set_deposition_binning(c, 0, origin, cell, nbin, data, t_end);
}

int arm_set_deposition_binning(arm_ctx* c, int type, const double* origin, const double* cell,
                               const int* nbin, const double* data, double t_end) {
  if (!c) return ARM_ERR_ARG;
  if (data && type != 0 && type != 1 && type != 7)
    return fail(c, ARM_ERR_UNSUPPORTED, "binning type %d: 0 (Cartesian), 1 or 7 (R-phi-z)", type);
  if (!data) {  // switch the source off
    if (c->d_dep) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipFree(c->d_dep));
    }
    c->d_dep = nullptr;
    c->K.dep = nullptr;
    return ARM_OK;
  }
  if (!origin || !cell || !nbin) return fail(c, ARM_ERR_ARG, "null origin, cell or nbin");
  size_t count = 1;
  for (int k = 0; k < 3; ++k) {
    if (nbin[k] < 1 || !(cell[k] > 0) || !std::isfinite(origin[k]))
      return fail(c, ARM_ERR_ARG, "bin %d: n = %d, width %g, origin %g", k, nbin[k], cell[k],
                  origin[k]);
    count *= (size_t)nbin[k];
  }
  if (count > ((size_t)1 << 31)) return fail(c, ARM_ERR_ARG, "map of %zu bins", count);
  // the 2D step reads the slice z = 0 (fluka_heat.f:488-493)
  if (!is3d(c) && !(origin[2] <= 0.0 && 0.0 <= origin[2] + nbin[2] * cell[2]))
    return fail(c, ARM_ERR_ARG, "z = 0 lies outside the binning's z range [%g, %g]", origin[2],
                origin[2] + nbin[2] * cell[2]);
  if (std::isnan(t_end)) return fail(c, ARM_ERR_ARG, "t_end is NaN");
  if (c->d_dep) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(c->d_dep));
    c->d_dep = nullptr;
    c->K.dep = nullptr;
  }
  void* p = nullptr;
  HIPCHK(c, hipMalloc(&p, count * sizeof(double)));
  c->d_dep = static_cast<double*>(p);
  HIPCHK(c, hipMemcpyAsync(p, data, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->K.dep = c->d_dep;
  for (int k = 0; k < 3; ++k) {
    c->K.dep_n[k] = nbin[k];
    c->K.dep_o[k] = origin[k];
    c->K.dep_d[k] = cell[k];
  }
  c->K.dep_type = type == 0 ? 0 : 1;
  c->dep_tend = t_end;
  return ARM_OK;
}

int arm_set_objective(arm_ctx* c, int kind) {
  if (!c) return ARM_ERR_ARG;
  if (kind != ARM_OBJ_PRESSURE && kind != ARM_OBJ_ENERGY)
    return fail(c, ARM_ERR_ARG, "objective kind %d", kind);
  c->objkind = kind;
  drop_adjoint(c);
  return ARM_OK;
}

int arm_rollout_param_grads(arm_ctx* c, double* out, int n) {
  if (!c || !out || n < 0) return ARM_ERR_ARG;
  for (int k = 0; k < n && k < ARM_NGRAD; ++k) out[k] = c->grads[k];
  return ARM_OK;
}

int arm_set_gather_path(arm_ctx* c, int tiled) {
  if (!c) return ARM_ERR_ARG;
  c->tiled = tiled != 0;
  return ARM_OK;
}

int arm_set_timing(arm_ctx* c, int on) {
  if (!c) return ARM_ERR_ARG;
  c->timing = on != 0;
  return ARM_OK;
}

}  // extern "C"

// =============================================================================
// arm_link_list_static: the cell keys, the counting sort and the cell ranges on
// caller device buffers, with the sort's workspace in a static device array
// (no allocation on any call).
// =============================================================================
namespace {

constexpr size_t al256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t ST_N = ARM_STATIC_MAX_PARTICLES;
constexpr size_t ST_C = ARM_STATIC_MAX_CELLS;
constexpr size_t ST_TILES = (ST_C + CS_TILE - 1) / CS_TILE;
static_assert(ST_TILES <= (size_t)CS_MAX_TILES, "static grid beyond the counting sort");
// regions: per-cell counters, tile sums, key by particle, (id, slot) entries
// grouped by cell, their cells, the outside count when the caller passes none
constexpr size_t ST_COUNT = 0;
constexpr size_t ST_TSUM = ST_COUNT + al256(4 * ST_C);
constexpr size_t ST_KEY = ST_TSUM + al256(4 * ST_TILES);
constexpr size_t ST_ENT = ST_KEY + al256(4 * ST_N);
constexpr size_t ST_ENTC = ST_ENT + al256(8 * ST_N);
constexpr size_t ST_OUT = ST_ENTC + al256(4 * ST_N);
constexpr size_t ST_BYTES = ST_OUT + 256;

}  // namespace

// One copy per GPU, held until the process ends (about 20 MB).  Host code
// never uses this name as an address: static_scratch() gives each device's.
static __device__ __align__(256) uint8_t scratch[ST_BYTES];

namespace {

constexpr int ST_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_dev[ST_MAX_DEVICES] = {};

// Address of the current device's copy of scratch, looked up on the first
// call on that device.
int static_scratch(uint8_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return ARM_ERR_HIP;
  if (dev < 0 || dev >= ST_MAX_DEVICES) return ARM_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return ARM_ERR_HIP;
    g_scratch_dev[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch_dev[dev];
  return ARM_OK;
}

// Keys of the particles of x(2, n) in id order (slot t = particle t), with
// the per-cell counts of the counting sort; outside counts the particles
// outside the grid (their key is that of cell 0, as in k_keys).
__global__ __launch_bounds__(CS_THREADS) void k_keys_xy(const double* __restrict__ x, int n,
                                                        KParams K, KeyOut ko,
                                                        int* __restrict__ outside) {
  const int t = blockIdx.x * CS_THREADS + threadIdx.x;
  const bool valid = t < n;
  uint32_t key = 0;
  if (valid) {
    int ix, iy;
    bool ok;
    key = (uint32_t)cell_of(K, x[2 * (size_t)t], x[2 * (size_t)t + 1], ix, iy, ok);
    if (!ok) atomicAdd(outside, 1);
  }
  emit_key(ko, valid, t, t, key);
}

}  // namespace

extern "C" int arm_link_list_static(const arm_params* prm, int n, const double* x,
                                    int32_t* key_sorted, int32_t* id_sorted,
                                    int32_t* cell_start, int32_t* outside, void* stream) {
  if (!prm || n < 0 || !cell_start || (n > 0 && (!x || !key_sorted || !id_sorted)))
    return ARM_ERR_ARG;
  if (prm->dim == 3) return ARM_ERR_UNSUPPORTED;
  if ((prm->dim != 0 && prm->dim != 2) || !(prm->cell > 0) || prm->nx < 1 || prm->ny < 1)
    return ARM_ERR_ARG;
  const long long ncell_ll = (long long)prm->nx * prm->ny;
  if (n > ARM_STATIC_MAX_PARTICLES || ncell_ll > ARM_STATIC_MAX_CELLS) return ARM_ERR_CAPACITY;
  // k_cs_tile_scan stores cell_start eight entries at a time in 16-byte words
  if ((reinterpret_cast<uintptr_t>(cell_start) & 15u) != 0u) return ARM_ERR_ARG;
  uint8_t* base = nullptr;
  const int rc = static_scratch(&base);
  if (rc) return rc;
  const int ncell = (int)ncell_ll;
  KParams K{};
  K.n = n;
  K.nx = prm->nx;
  K.ny = prm->ny;
  K.ncell = ncell;
  K.ox = prm->ox;
  K.oy = prm->oy;
  K.cell = prm->cell;
  int* count = reinterpret_cast<int*>(base + ST_COUNT);
  int* tsum = reinterpret_cast<int*>(base + ST_TSUM);
  uint32_t* key = reinterpret_cast<uint32_t*>(base + ST_KEY);
  uint32_t* ent_id = reinterpret_cast<uint32_t*>(base + ST_ENT);
  const Ent ent{ent_id, ent_id + ST_N};
  uint32_t* entc = reinterpret_cast<uint32_t*>(base + ST_ENTC);
  int* out = outside ? outside : reinterpret_cast<int*>(base + ST_OUT);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int tiles = (ncell + CS_TILE - 1) / CS_TILE;
  const int nb = (n + CS_THREADS - 1) / CS_THREADS;
  // fill kernels, not memsets: a caller may capture this call in a graph
  hipLaunchKernelGGL(k_cs_zero, dim3(std::min(1024, (ncell + CS_THREADS - 1) / CS_THREADS)),
                     dim3(CS_THREADS), 0, s, count, ncell);
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(BLOCK), 0, s, reinterpret_cast<uint32_t*>(out),
                     (size_t)1);
  if (n > 0)
    hipLaunchKernelGGL(k_keys_xy, dim3(nb), dim3(CS_THREADS), 0, s, x, n, K,
                       KeyOut{nullptr, nullptr, key, count}, out);
  hipLaunchKernelGGL(k_cs_tile_sum, dim3(tiles), dim3(CS_THREADS), 0, s, count, ncell, tsum,
                     Halt{});
  hipLaunchKernelGGL(k_cs_tile_scan, dim3(tiles), dim3(CS_THREADS), 0, s, count, ncell, n, tsum,
                     reinterpret_cast<int*>(cell_start), Halt{});
  if (n > 0) {
    hipLaunchKernelGGL(k_cs_scatter, dim3(nb), dim3(CS_THREADS), 0, s, key,
                       static_cast<const int32_t*>(nullptr), n,
                       reinterpret_cast<const int*>(cell_start), count, ent, entc, Halt{});
    hipLaunchKernelGGL(k_cs_rank, dim3(nb), dim3(CS_THREADS), 0, s, ent, entc, n,
                       reinterpret_cast<const int*>(cell_start),
                       reinterpret_cast<uint32_t*>(key_sorted),
                       reinterpret_cast<uint32_t*>(id_sorted), -1, Halt{});
  }
  return hipGetLastError() == hipSuccess ? ARM_OK : ARM_ERR_HIP;
}
