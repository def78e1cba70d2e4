// extern "C" entry points of libsymgw (include/symgw.h).
//
// A context is the device residency of one system: it plays the role of
// cugw_routine_block<prec> (SymmetryMBPT/src/cuda/cugw_routines_block.cu:6-91)
// but lives across GW iterations, owns one HIP stream and keeps G, P, Sigma
// and V resident.  The host-side W and Dyson solves of the reference
// (cugw_solver_block.cpp:385-439, sc_type_block.cpp:6-87) run on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/symgw.h"
#include "common.h"
#include "kernels.h"
#include "layout.h"

using namespace symgw;

namespace {

thread_local std::string g_create_error;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct symgw_ctx {
  Layout L;
  WorkLists W;
  int device = 0;
  int prec = SYMGW_PREC_FP64;
  int eb = 16;  // bytes per complex element in the contraction precision
  hipStream_t st = nullptr;
  std::string err;
  std::vector<DevBuf> owned;

  // resident data
  void *V = nullptr;                   // flat V, contraction precision
  cplx<double> *G_ibz = nullptr;       // ts_kij, FP64
  void *G_full = nullptr;              // k_ts_ij full BZ, contraction precision
  cplx<double> *P = nullptr;           // q_t_QP iBZ, FP64 (P0, then P)
  void *P_full = nullptr;              // q_t_QP full BZ, contraction precision
  cplx<double> *Sig = nullptr;         // k_ts_ij iBZ, FP64
  cplx<double> *Pw = nullptr;          // [q_ir][n][qslice]
  cplx<double> *SGw = nullptr;         // [k_ir][w][s][kslice] (Sigma(iw), then G(iw))
  cplx<double> *F = nullptr, *S = nullptr;
  cplx<double> *orep = nullptr, *auxrep = nullptr;
  cplx<double> *Tnt = nullptr, *Ttn = nullptr, *Tnt_BF = nullptr, *Ttn_FB = nullptr, *iwn = nullptr;
  cplx<double> *w_scratch = nullptr, *d_scratch = nullptr;
  double *kw = nullptr, *nel_dev = nullptr;
  // shard: the P0 / Sigma sweeps cover iBZ q / k indices [shard_first, shard_end)
  int shard_first = 0, shard_end = 0;
  // streamed V: generator parameters and the per-batch generator blocks
  bool stream = false;
  uint64_t gen_seed = 0;
  double gen_scale = 0.0;
  double *d_decay = nullptr;
  int n_decay = 0;
  GenBlk *d_p0_gen = nullptr, *d_sig_gen = nullptr;
  std::vector<int> p0_gen_begin, sig_gen_begin;  // [ink + 1] prefixes into the generator block lists
  // mixed mode: INT8 slice copies of V (Sigma's blocks) and of P (full BZ)
  int8_t *V8 = nullptr, *P8 = nullptr;
  int *VE = nullptr, *PE = nullptr;
  Split8Blk *d_v8_split = nullptr, *d_p8_split = nullptr;
  int64_t *d_v8_prefix = nullptr, *d_p8_prefix = nullptr;
  int64_t v8_rows = 0, p8_rows = 0;
  std::vector<int64_t> v8_batch_rows;  // streamed V: split rows of each Sigma batch
  Plan plan;                           // planner rules of the work lists

  // device tables
  P0Item *d_p0_items = nullptr;
  P0Blk *d_p0_blks = nullptr;
  SigItem *d_sig_items[kNumSigDepths] = {};  // slabs of kSigDepths[d] times
  SigBlk *d_sig_blks = nullptr;
  // split P0 / Sigma items, and the partial sums of their parts (one buffer,
  // the P0 and Sigma sweeps run one after the other on the context stream)
  P0Split *d_p0_splits = nullptr;
  SigSplit *d_sig_splits = nullptr;
  cplx<double> *part = nullptr;
  int64_t part_cap = 0;
  RotBlk *d_g_rot = nullptr, *d_p_rot = nullptr;
  SolveBlk *d_w_blocks = nullptr, *d_d_blocks = nullptr;
  int64_t *d_g_src_stride = nullptr, *d_g_src_slice = nullptr, *d_g_dst_slice = nullptr, *d_g_dst_size = nullptr;
  int64_t *d_p_src_stride = nullptr, *d_p_src_slice = nullptr, *d_p_dst_slice = nullptr, *d_p_dst_size = nullptr;
  int64_t *d_w_off = nullptr, *d_qslice_irre = nullptr, *d_qoff_irre = nullptr;
  int64_t *d_d_off = nullptr, *d_kslice_irre = nullptr, *d_kao_off_irre = nullptr;
  TBatch *d_tb_wf = nullptr, *d_tb_wb = nullptr, *d_tb_df = nullptr, *d_tb_db = nullptr;
  int max_qslice = 0, max_kslice = 0;

  bool have_V = false, have_G = false, have_reps = false, have_T = false, have_F = false, have_P = false,
       have_Sig = false;
  bool p_rotated = false;  // P_full holds the rotation of the current P
  bool timing = false;
  double last_ms[5] = {0, 0, 0, 0, 0};
  hipEvent_t ev[6] = {};
  // one hipGraph per phase (P0, W, Sigma, Dyson), replayed back to back on the
  // context stream; with timing on, events sit between the replays
  hipGraphExec_t graph_exec[4] = {};
  double graph_mu = 0.0;
};

namespace {

int fail(symgw_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

// Runs the body of an extern "C" entry point so that no C++ exception crosses
// the ABI: a failed host allocation becomes SYMGW_ERR_NOMEM, anything else
// SYMGW_ERR_ARG, with the message on the context (or, without one, in the
// thread's create-error string).
template <typename F>
int guarded(symgw_ctx *c, F &&body) noexcept {
  auto report = [c](int code, const char *what, const char *detail) noexcept {
    try {
      std::string msg = std::string(what) + detail;
      if (c) c->err = msg;
      else g_create_error = msg;
    } catch (...) {
    }
    return code;
  };
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return report(SYMGW_ERR_NOMEM, "host memory allocation failed", "");
  } catch (const std::exception &e) {
    return report(SYMGW_ERR_ARG, "internal error: ", e.what());
  } catch (...) {
    return report(SYMGW_ERR_ARG, "internal error", "");
  }
}

#define HIPCHK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, SYMGW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Allocation, release and graph capture are serialised over all contexts of
// the process: hipMalloc / hipFree must not run on one thread while another
// thread captures a graph.  Copies and memsets go through the context's own
// non-blocking stream; null-stream operations would have to wait on (and so
// break) a capture running on another context's thread.
std::mutex g_dev_mu;

hipError_t copy_sync(symgw_ctx *c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  return e;
}

hipError_t dev_malloc(void **p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  return hipMalloc(p, bytes);
}

void dev_free(void *p) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  (void)hipFree(p);
}

// Context buffers come from the device's stream-ordered pool, allocated and
// freed on the context's own stream: hipFree waits for the whole device, so a
// context set up or released while others run (serving) would stall them all;
// hipFreeAsync waits for nothing but its stream.  The pool is trimmed after a
// release so the memory goes back to the device.
hipError_t ctx_malloc(symgw_ctx *c, void **p, size_t bytes) { return hipMallocAsync(p, bytes, c->st); }

void ctx_free(symgw_ctx *c, void *p) { (void)hipFreeAsync(p, c->st); }

void trim_pool(int device) {
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) (void)hipMemPoolTrimTo(pool, 0);
}

void drop_graphs(symgw_ctx *c);

// Release every buffer of a context, its graphs, events and stream.
void release_ctx(symgw_ctx *c) {
  for (auto &b : c->owned) ctx_free(c, b.p);
  c->owned.clear();
  (void)hipStreamSynchronize(c->st);
  {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    drop_graphs(c);
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->st);
  }
  trim_pool(c->device);
}

template <typename X>
int dalloc(symgw_ctx *c, X **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  void *q = nullptr;
  hipError_t e = ctx_malloc(c, &q, count * sizeof(X));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, SYMGW_ERR_NOMEM, "hipMallocAsync of " + std::to_string(count * sizeof(X)) + " bytes failed");
  }
  c->owned.push_back(DevBuf{q, count * sizeof(X)});
  *p = static_cast<X *>(q);
  return SYMGW_OK;
}
int dalloc_bytes(symgw_ctx *c, void **p, size_t bytes) {
  char *q = nullptr;
  int r = dalloc(c, &q, bytes);
  *p = q;
  return r;
}

template <typename X>
int upload(symgw_ctx *c, X **dst, const std::vector<X> &v) {
  int r = dalloc(c, dst, v.size());
  if (r) return r;
  if (!v.empty()) HIPCHK(c, copy_sync(c, *dst, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
  return SYMGW_OK;
}

int64_t footprint(const Layout &L, const WorkLists &W, int eb) {
  int64_t b = 0;
  b += (W.stream ? W.batch_V_max : L.flat_V) * eb;  // resident V, or the largest streamed batch
  b += (int64_t)L.nts * L.ns * L.flat_ao * 16 * 2;       // G_ibz, Sig
  b += (int64_t)L.nts * L.ns * L.flat_ao_full * eb;      // G_full
  b += (int64_t)L.nts * L.flat_aux * 16;                 // P
  b += (int64_t)L.nts * L.flat_aux_full * eb;            // P_full
  b += (int64_t)L.nw_b * L.flat_aux * 16 + (int64_t)L.nw * L.ns * L.flat_ao * 16;
  b += (int64_t)L.nw_b * W.w_scr_group * 16 + (int64_t)L.nw * L.ns * W.d_scr_group * 16;
  b += (int64_t)(W.p0_blks.size() * sizeof(P0Blk) + W.sig_blks.size() * sizeof(SigBlk));
  // split items: their tables and the partial sums of their parts
  b += std::max(W.p0_part_elems, W.sig_part_elems) * 16 +
       (int64_t)(W.p0_splits.size() * sizeof(P0Split) + W.sig_splits.size() * sizeof(SigSplit));
  // mixed mode: INT8 slices, exponents and split tables
  b += W.v8_bytes + W.p8_bytes + 4 * (W.ve_count + W.pe_count);
  b += (int64_t)(W.v8_split.size() + W.p8_split.size()) * (sizeof(Split8Blk) + 8);
  return b;
}

// precision argument: a SYMGW_PREC_* mode, optionally with SYMGW_FLAG_STREAM_V
int prec_mode(int p) { return p & 0xFF; }
bool prec_stream(int p) { return (p & SYMGW_FLAG_STREAM_V) != 0; }
bool valid_precision(int p) {
  const int m = prec_mode(p);
  return (p & ~(0xFF | SYMGW_FLAG_STREAM_V)) == 0 &&
         (m == SYMGW_PREC_FP64 || m == SYMGW_PREC_FP32 || m == SYMGW_PREC_MIXED);
}
// bytes of one complex element of V, G and P in the contractions
int precision_bytes(int p) { return prec_mode(p) == SYMGW_PREC_FP32 ? 8 : 16; }

// Generator blocks of the pairs of every streamed batch (V offsets inside the
// batch buffer, counters of the resident layout so the values are the same).
void batch_gen_blocks(const Layout &L, const std::vector<std::pair<int, int64_t>> &pairs,
                      const std::vector<int> &pair_begin, std::vector<GenBlk> &out, std::vector<int> &begin) {
  begin.assign(1, 0);
  for (size_t bt = 0; bt + 1 < pair_begin.size(); ++bt) {
    for (int i = pair_begin[bt]; i < pair_begin[bt + 1]; ++i) {
      const int p = pairs[i].first, k1 = p / L.nk, k2 = p % L.nk;
      for (size_t b = 0; b < L.V_irreps[p].size(); ++b) {
        const auto &t = L.V_irreps[p][b];
        GenBlk g;
        g.v_off = pairs[i].second + L.V_offsets[p][b];
        g.local_off = L.V_offsets[p][b];
        g.q = L.V_qsize[p][b];
        g.k1 = L.ao_sizes[k1][t[1]];
        g.k2 = L.ao_sizes[k2][t[2]];
        g.a_base = L.ao_block_offsets[k1][t[1]];
        g.b_base = L.ao_block_offsets[k2][t[2]];
        g.pair = p;
        out.push_back(g);
      }
    }
    begin.push_back((int)out.size());
  }
}

// Streamed V: generate batch `bt` of a contraction into the batch buffer.
template <typename T>
hipError_t gen_batch(symgw_ctx *c, const GenBlk *d_gen, const std::vector<int> &begin, int bt) {
  return launch_genv<T>(d_gen + begin[bt], begin[bt + 1] - begin[bt], c->gen_seed, c->gen_scale, c->d_decay,
                        c->n_decay, static_cast<cplx<T> *>(c->V), c->st);
}

// INT8 slices of V (after it is set) and of P_full (after each rotation).
int split_V8(symgw_ctx *c) {
  if (c->prec != SYMGW_PREC_MIXED || c->stream) return SYMGW_OK;  // streamed: split per batch
  hipError_t e = launch_split8(c->d_v8_split, (int)c->W.v8_split.size(), c->d_v8_prefix, c->v8_rows,
                               static_cast<const cplx<double> *>(c->V), c->V8, c->VE, c->st);
  if (e != hipSuccess) return fail(c, SYMGW_ERR_HIP, std::string("V split: ") + hipGetErrorString(e));
  return SYMGW_OK;
}
int split_P8(symgw_ctx *c) {
  if (c->prec != SYMGW_PREC_MIXED) return SYMGW_OK;
  hipError_t e = launch_split8(c->d_p8_split, (int)c->W.p8_split.size(), c->d_p8_prefix, c->p8_rows,
                               static_cast<const cplx<double> *>(c->P_full), c->P8, c->PE, c->st);
  if (e != hipSuccess) return fail(c, SYMGW_ERR_HIP, std::string("P split: ") + hipGetErrorString(e));
  return SYMGW_OK;
}

template <typename T>
int run_rotate_G(symgw_ctx *c) {
  const Layout &L = c->L;
  hipError_t e = launch_rotate<T>(c->d_g_rot, (int)c->W.g_rot.size(), L.nts * L.ns, c->G_ibz, c->d_g_src_stride,
                                  c->d_g_src_slice, c->orep, static_cast<cplx<T> *>(c->G_full), c->d_g_dst_slice,
                                  c->d_g_dst_size, c->st);
  if (e != hipSuccess) return fail(c, SYMGW_ERR_HIP, std::string("rotate G: ") + hipGetErrorString(e));
  return SYMGW_OK;
}

int rotate_G(symgw_ctx *c) {
  if (!c->have_reps && c->L.flat_orep > 0) return fail(c, SYMGW_ERR_STATE, "rotation matrices not set");
  return c->eb == 16 ? run_rotate_G<double>(c) : run_rotate_G<float>(c);
}

template <typename T>
int run_rotate_P(symgw_ctx *c) {
  const Layout &L = c->L;
  hipError_t e = launch_rotate<T>(c->d_p_rot, (int)c->W.p_rot.size(), L.nts, c->P, c->d_p_src_stride,
                                  c->d_p_src_slice, c->auxrep, static_cast<cplx<T> *>(c->P_full), c->d_p_dst_slice,
                                  c->d_p_dst_size, c->st);
  if (e != hipSuccess) return fail(c, SYMGW_ERR_HIP, std::string("rotate P: ") + hipGetErrorString(e));
  return SYMGW_OK;
}

template <typename T>
int run_p0(symgw_ctx *c) {
  const Layout &L = c->L;
  const double pref = L.ns == 1 ? -2.0 : -1.0;  // cugw_qkpt_block.cu:210
  hipError_t e = hipSuccess;
  const auto &ib = c->W.p0_batch_items;
  if (!c->stream) {
    // the items of the shard's iBZ q, in one launch
    e = launch_p0<T>(c->d_p0_items + ib[c->shard_first], c->d_p0_blks, ib[c->shard_end] - ib[c->shard_first], L.nts,
                     L.ns, static_cast<const cplx<T> *>(c->V), static_cast<const cplx<T> *>(c->G_full), c->P, c->part,
                     pref, c->W.p0_lds, c->st, c->W.p0_mode);
  } else {
    // one batch per group of iBZ q: its V pairs generated, then the items of
    // its points inside the shard contracted in one launch
    const auto &gf = c->W.p0_group_first;
    for (size_t g = 0; g + 1 < gf.size() && e == hipSuccess; ++g) {
      const int a = std::max(gf[g], c->shard_first), b = std::min(gf[g + 1], c->shard_end);
      if (a >= b) continue;
      e = gen_batch<T>(c, c->d_p0_gen, c->p0_gen_begin, (int)g);
      if (e == hipSuccess)
        e = launch_p0<T>(c->d_p0_items + ib[a], c->d_p0_blks, ib[b] - ib[a], L.nts, L.ns,
                         static_cast<const cplx<T> *>(c->V), static_cast<const cplx<T> *>(c->G_full), c->P, c->part,
                         pref, c->W.p0_lds, c->st, c->W.p0_mode);
    }
  }
  // the split items' parts, added in order
  if (e == hipSuccess)
    e = launch_p0_reduce(c->d_p0_splits, (int)c->W.p0_splits.size(), c->shard_first, c->shard_end, c->part, c->P,
                         pref, L.nts, c->st);
  if (e != hipSuccess) return fail(c, SYMGW_ERR_HIP, std::string("P0 kernel: ") + hipGetErrorString(e));
  return SYMGW_OK;
}

template <typename T>
int run_sigma(symgw_ctx *c) {
  const Layout &L = c->L;
  hipError_t e = hipSuccess;
  Z8Ops z8;
  z8.P8 = c->P8;
  z8.V8 = c->V8;
  z8.PE = c->PE;
  z8.VE = c->VE;
  const Z8Ops *pz8 = c->prec == SYMGW_PREC_MIXED ? &z8 : nullptr;
  if (!c->stream) {
    // the items of the shard's iBZ k, one launch per slab depth
    for (int d = 0; d < kNumSigDepths && e == hipSuccess; ++d) {
      const auto &ib = c->W.sig_batch_items[d];
      e = launch_sigma<T>(c->d_sig_items[d] + ib[c->shard_first], c->d_sig_blks, ib[c->shard_end] - ib[c->shard_first],
                          L.nts, L.ns, static_cast<const cplx<T> *>(c->V), static_cast<const cplx<T> *>(c->G_full),
                          static_cast<const cplx<T> *>(c->P_full), c->Sig, c->part, -1.0, c->W.sig_lds[d],
                          c->st, kSigDepths[d], pz8);
    }
  } else {
    // one batch per group of iBZ k: its V pairs generated (and split, in the
    // mixed mode), then the items of its points inside the shard, per slab depth
    const auto &gf = c->W.sig_group_first;
    for (size_t g = 0; g + 1 < gf.size() && e == hipSuccess; ++g) {
      const int a = std::max(gf[g], c->shard_first), b = std::min(gf[g + 1], c->shard_end);
      if (a >= b) continue;
      e = gen_batch<T>(c, c->d_sig_gen, c->sig_gen_begin, (int)g);
      if (e == hipSuccess && pz8) {
        const int i0 = c->W.v8_batch_begin[g];
        e = launch_split8(c->d_v8_split + i0, c->W.v8_batch_begin[g + 1] - i0, c->d_v8_prefix + i0,
                          c->v8_batch_rows[g], static_cast<const cplx<double> *>(c->V), c->V8, c->VE, c->st);
      }
      for (int d = 0; d < kNumSigDepths && e == hipSuccess; ++d) {
        const auto &ib = c->W.sig_batch_items[d];
        if (ib[b] == ib[a]) continue;
        e = launch_sigma<T>(c->d_sig_items[d] + ib[a], c->d_sig_blks, ib[b] - ib[a], L.nts, L.ns,
                            static_cast<const cplx<T> *>(c->V), static_cast<const cplx<T> *>(c->G_full),
                            static_cast<const cplx<T> *>(c->P_full), c->Sig, c->part, -1.0,
                            c->W.sig_lds[d], c->st, kSigDepths[d], pz8);
      }
    }
  }
  // the split items' parts, added in order (every launch above is on the same stream)
  if (e == hipSuccess)
    e = launch_sig_reduce(c->d_sig_splits, (int)c->W.sig_splits.size(), c->shard_first, c->shard_end, c->part,
                          c->Sig, -1.0, L.nts, c->st);
  if (e != hipSuccess) return fail(c, SYMGW_ERR_HIP, std::string("Sigma kernel: ") + hipGetErrorString(e));
  return SYMGW_OK;
}

// P (iBZ q) -> P_full (every q of the full BZ).
int rotate_P(symgw_ctx *c) {
  if (!c->have_reps && c->L.flat_auxrep > 0) return fail(c, SYMGW_ERR_STATE, "rotation matrices not set");
  int r = c->eb == 16 ? run_rotate_P<double>(c) : run_rotate_P<float>(c);
  if (r) return r;
  if ((r = split_P8(c))) return r;
  c->p_rotated = true;
  return SYMGW_OK;
}

int phase_P0(symgw_ctx *c) {
  if (!c->have_V) return fail(c, SYMGW_ERR_STATE, "V not set");
  if (!c->have_G) return fail(c, SYMGW_ERR_STATE, "G not uploaded");
  const Layout &L = c->L;
  int r = c->eb == 16 ? run_p0<double>(c) : run_p0<float>(c);
  if (r) return r;
  (void)L;
  c->p_rotated = false;
  return SYMGW_OK;
}

// The W phase with the solver scratch at `scratch`, which holds `per_launch`
// frequency groups (the solves run in launches of at most that many
// frequencies).
int phase_W_with(symgw_ctx *c, cplx<double> *scratch, int64_t per_launch) {
  if (!c->have_T) return fail(c, SYMGW_ERR_STATE, "transforms not set");
  const Layout &L = c->L;
  const int KT = L.nts - 2;
  // P0(iW_n) = Tnt_BF * P0(t interior), mirrored, / nk (cugw_solver_block.cpp:395-400)
  HIPCHK(c, launch_tgemm(c->d_tb_wf, L.ink, c->max_qslice, L.nw_b, KT, 1, c->Tnt_BF, c->P, c->Pw, 1, 1, L.nts,
                         1.0 / L.nk, 0, 0, c->st));
  for (int n0 = 0; n0 < L.nw_b; n0 += (int)per_launch)
    HIPCHK(c, launch_wsolve(c->d_w_blocks, (int)c->W.w_blocks.size(), n0, (int)std::min<int64_t>(per_launch, L.nw_b - n0),
                            c->Pw, c->d_w_off, c->d_qslice_irre, scratch, c->W.w_scr_group, c->st));
  // P(t) = Ttn_FB * Pt(iW_n) (cugw_solver_block.cpp:428)
  HIPCHK(c, launch_tgemm(c->d_tb_wb, L.ink, c->max_qslice, L.nts, L.nw_b, 1, c->Ttn_FB, c->Pw, c->P, 0, 0, L.nts,
                         1.0, 0, 0, c->st));
  c->have_P = true;
  // P of every full-BZ q for the Sigma contraction (cugw_qkpt_block.cu:287-340)
  return rotate_P(c);
}

int phase_W(symgw_ctx *c) { return phase_W_with(c, c->w_scratch, std::max(1, c->L.nw_b)); }

int phase_Sigma(symgw_ctx *c) {
  if (!c->have_V || !c->have_G) return fail(c, SYMGW_ERR_STATE, "V or G missing");
  if (!c->have_P) return fail(c, SYMGW_ERR_STATE, "P not available");
  int r = c->p_rotated ? SYMGW_OK : rotate_P(c);
  if (r) return r;
  r = c->eb == 16 ? run_sigma<double>(c) : run_sigma<float>(c);
  if (r) return r;
  c->have_Sig = true;
  return SYMGW_OK;
}

int phase_Dyson(symgw_ctx *c, double mu, double sigma_scale) {
  if (!c->have_T || !c->have_F) return fail(c, SYMGW_ERR_STATE, "transforms or Fock not set");
  const Layout &L = c->L;
  const int KT = L.nts - 2;
  if (c->have_Sig) {
    // Sigma(iw) = Tnt * Sigma(t interior) (sc_type_block.cpp:150)
    HIPCHK(c, launch_tgemm(c->d_tb_df, L.ink, c->max_kslice * L.ns, L.nw, KT, 1, c->Tnt, c->Sig, c->SGw, 0, 1,
                           L.nts, 1.0, 0, 0, c->st));
  } else {
    sigma_scale = 0.0;  // Sigma = 0: Dyson of the bare Fock matrix (sc_loop init)
  }
  HIPCHK(c, launch_dyson(c->d_d_blocks, (int)c->W.d_blocks.size(), L.nw, L.ns, c->SGw, c->SGw, c->d_d_off,
                         c->d_kslice_irre, c->d_kao_off_irre, c->F, c->S, L.flat_ao, c->iwn, mu, sigma_scale,
                         c->d_scratch, c->W.d_scr_group, c->st));
  // G(t) = Ttn * G(iw) -> ts_kij (sc_type_block.cpp:171)
  HIPCHK(c, launch_tgemm(c->d_tb_db, L.ink, c->max_kslice * L.ns, L.nts, L.nw, L.ns, c->Ttn, c->SGw, c->G_ibz, 0, 0,
                         L.nts, 1.0, 0, 0, c->st));
  c->have_G = true;
  return rotate_G(c);
}

void ev_record(symgw_ctx *c, int i) {
  if (c->timing) (void)hipEventRecord(c->ev[i], c->st);
}

int run_phase(symgw_ctx *c, int phase, double mu) {
  switch (phase) {
    case 0: return phase_P0(c);
    case 1: return phase_W(c);
    case 2: return phase_Sigma(c);
    default: return phase_Dyson(c, mu, 1.0 / c->L.nk);
  }
}

int iteration_body(symgw_ctx *c, double mu) {
  int r;
  ev_record(c, 0);
  for (int ph = 0; ph < 4; ++ph) {
    if ((r = run_phase(c, ph, mu))) return r;
    ev_record(c, ph + 1);
  }
  return SYMGW_OK;
}

void drop_graphs(symgw_ctx *c) {
  for (auto &g : c->graph_exec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

}  // namespace

namespace {
template <typename T>
int time_mfma_peak(hipStream_t st, hipEvent_t e0, hipEvent_t e1, void *out, int n_wg, int iters, double *tflops) {
  if (launch_mfma_peak<T>(out, n_wg, 16, st) != hipSuccess) return SYMGW_ERR_HIP;  // warm-up
  if (hipEventRecord(e0, st) != hipSuccess || launch_mfma_peak<T>(out, n_wg, iters, st) != hipSuccess ||
      hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
    return SYMGW_ERR_HIP;
  float ms = 0;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0) return SYMGW_ERR_HIP;
  // waves x iterations x 2 tiles x 4 MFMAs x 2*16*16*4 flops
  const double flops = (double)n_wg * (mfma_peak_threads() / 64) * iters * 2 * 4 * 2048.0;
  *tflops = flops / (ms * 1e-3) / 1e12;
  return SYMGW_OK;
}
}  // namespace

// Scratch of symgw_solve_W_noalloc: a file-scope device array, one copy per
// GPU the library runs on, held until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the current device
// (kept per device in g_noalloc_scratch); the array's name in host code is a
// stand-in, not a device address.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
// Serialises the allocation-free calls of this process (they share each GPU's
// array) and guards the per-device addresses.
std::mutex g_noalloc_mu;
void *g_noalloc_scratch[kNoAllocMaxDevices] = {};
}  // namespace

// Replace a device table: the old buffer (if any) is released on the
// context's stream, the new one uploaded.
template <typename X>
static int reupload(symgw_ctx *c, X **dst, const std::vector<X> &v) {
  if (*dst) {
    void *old = *dst;
    for (size_t i = 0; i < c->owned.size(); ++i)
      if (c->owned[i].p == old) {
        c->owned.erase(c->owned.begin() + (long)i);
        break;
      }
    ctx_free(c, old);
    *dst = nullptr;
  }
  return upload(c, dst, v);
}

// The device tables that follow from the plan: P0 and Sigma items and
// blocks, and in the mixed mode the INT8 split tables and their row prefixes.
static int upload_plan_tables(symgw_ctx *c) {
  int r = reupload(c, &c->d_p0_items, c->W.p0_items);
  r = r ? r : reupload(c, &c->d_p0_blks, c->W.p0_blks);
  for (int d = 0; d < kNumSigDepths; ++d) r = r ? r : reupload(c, &c->d_sig_items[d], c->W.sig_items[d]);
  r = r ? r : reupload(c, &c->d_sig_blks, c->W.sig_blks);
  r = r ? r : reupload(c, &c->d_p0_splits, c->W.p0_splits);
  r = r ? r : reupload(c, &c->d_sig_splits, c->W.sig_splits);
  const int64_t part_need = std::max(c->W.p0_part_elems, c->W.sig_part_elems);
  if (!r && part_need > c->part_cap) {
    // the partial sums of split items: a plan that splits more gets a larger buffer
    if (c->part) {
      for (size_t i = 0; i < c->owned.size(); ++i)
        if (c->owned[i].p == c->part) {
          c->owned.erase(c->owned.begin() + (long)i);
          break;
        }
      ctx_free(c, c->part);
      c->part = nullptr;
    }
    r = dalloc(c, &c->part, (size_t)part_need);
    if (!r) c->part_cap = part_need;
  }
  if (r || c->prec != SYMGW_PREC_MIXED) return r;
  auto prefix = [](const std::vector<Split8Blk> &v, int64_t &tot) {
    std::vector<int64_t> p(v.size() + 1, 0);
    for (size_t i = 0; i < v.size(); ++i) p[i + 1] = p[i] + (int64_t)v[i].nsub * v[i].nrows;
    tot = p.back();
    return p;
  };
  r = reupload(c, &c->d_v8_split, c->W.v8_split);
  r = r ? r : reupload(c, &c->d_p8_split, c->W.p8_split);
  if (!c->stream) {
    r = r ? r : reupload(c, &c->d_v8_prefix, prefix(c->W.v8_split, c->v8_rows));
  } else {
    // row prefixes restarted at every batch, so each batch's split launch
    // reads its own slice of the table
    const auto &vb = c->W.v8_batch_begin;
    std::vector<int64_t> pre(c->W.v8_split.size() + 1, 0);
    c->v8_batch_rows.assign(vb.empty() ? 0 : vb.size() - 1, 0);
    for (size_t bt = 0; bt + 1 < vb.size(); ++bt) {
      int64_t acc = 0;
      for (int i = vb[bt]; i < vb[bt + 1]; ++i) {
        pre[i] = acc;
        acc += (int64_t)c->W.v8_split[i].nsub * c->W.v8_split[i].nrows;
      }
      c->v8_batch_rows[bt] = acc;
    }
    r = r ? r : reupload(c, &c->d_v8_prefix, pre);
  }
  r = r ? r : reupload(c, &c->d_p8_prefix, prefix(c->W.p8_split, c->p8_rows));
  return r;
}

static Plan plan_of(const symgw_plan &p) {
  Plan q;
  q.sig_chunk_cost = p.sig_chunk_cost;
  q.sig_rows = p.sig_rows;
  q.sig_max_times = p.sig_max_times;
  q.p0_chunk_cost = p.p0_chunk_cost;
  q.p0_band_cap = p.p0_band_cap;
  return q;
}

extern "C" {

const char *symgw_create_error(void) { return g_create_error.c_str(); }

const char *symgw_last_error(const symgw_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int symgw_layout_query(const symgw_layout_desc *desc, symgw_layout_info *info) {
  return guarded(nullptr, [&]() -> int {
    if (!info) return SYMGW_ERR_ARG;
    Layout L;
    std::string err;
    if (!build_layout(desc, L, err)) {
      g_create_error = err;
      return SYMGW_ERR_ARG;
    }
    WorkLists W;
    if (!build_worklists(L, 16, W, err)) {
      g_create_error = err;
      return SYMGW_ERR_ARG;
    }
    info->flat_ao_size = L.flat_ao;
    info->flat_ao_size_full = L.flat_ao_full;
    info->flat_aux_size = L.flat_aux;
    info->flat_aux_size_full = L.flat_aux_full;
    info->flat_V_size = L.flat_V;
    info->flat_orep_size = L.flat_orep;
    info->flat_auxrep_size = L.flat_auxrep;
    info->n_P0_blocks = W.n_p0_contractions;
    info->n_Sigma_blocks = W.n_sig_contractions;
    info->flop_P0 = W.flop_p0;
    info->flop_Sigma = W.flop_sig;
    info->device_bytes = footprint(L, W, 16);
    return SYMGW_OK;
  });
}

int symgw_layout_footprint(const symgw_layout_desc *desc, int precision, int64_t *bytes) {
  return guarded(nullptr, [&]() -> int {
    if (!desc || !bytes) return SYMGW_ERR_ARG;
    if (!valid_precision(precision)) {
      g_create_error = "unknown precision mode";
      return SYMGW_ERR_ARG;
    }
    const int eb = precision_bytes(precision);
    Layout L;
    WorkLists W;
    std::string err;
    if (!build_layout(desc, L, err) ||
        !build_worklists(L, eb, W, err, prec_mode(precision) == SYMGW_PREC_MIXED, prec_stream(precision))) {
      g_create_error = err;
      return SYMGW_ERR_ARG;
    }
    *bytes = footprint(L, W, eb);
    return SYMGW_OK;
  });
}

int symgw_device_memory(int device, int64_t *free_bytes, int64_t *total_bytes) {
  return guarded(nullptr, [&]() -> int {
    if (!free_bytes || !total_bytes) return SYMGW_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      g_create_error = "no HIP device";
      return SYMGW_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) {
      g_create_error = "device index out of range";
      return SYMGW_ERR_ARG;
    }
    size_t fr = 0, tot = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
      g_create_error = "hipMemGetInfo failed";
      return SYMGW_ERR_HIP;
    }
    *free_bytes = (int64_t)fr;
    *total_bytes = (int64_t)tot;
    return SYMGW_OK;
  });
}

int symgw_measure_peaks(int device, double *out) {
  return guarded(nullptr, [&]() -> int {
    if (!out) return SYMGW_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      g_create_error = "no HIP device";
      return SYMGW_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) {
      g_create_error = "device index out of range";
      return SYMGW_ERR_ARG;
    }
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
      g_create_error = "device query failed";
      return SYMGW_ERR_HIP;
    }
    const int n_wg = 4 * prop.multiProcessorCount;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      g_create_error = "hipMemGetInfo failed";
      return SYMGW_ERR_HIP;
    }
    // two copy buffers of up to 1 GiB each, at most a quarter of the free memory
    const int64_t cbytes = std::min<int64_t>(int64_t(1) << 30, (int64_t)(fr / 4)) & ~int64_t(4095);
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void *res = nullptr, *src = nullptr, *dst = nullptr;
    int r = SYMGW_OK;
    std::string what;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
      r = SYMGW_ERR_HIP;
      what = "stream/event creation failed";
    }
    if (!r && (dev_malloc(&res, (size_t)n_wg * mfma_peak_threads() * sizeof(double)) != hipSuccess ||
               dev_malloc(&src, (size_t)cbytes) != hipSuccess || dev_malloc(&dst, (size_t)cbytes) != hipSuccess)) {
      r = SYMGW_ERR_NOMEM;
      what = "peak benchmark buffers";
    }
    // best of three runs of each (the first ones also bring the clocks up)
    out[0] = out[1] = out[2] = 0.0;
    for (int rep = 0; rep < 3 && !r; ++rep) {
      double t64 = 0, t32 = 0;
      if ((r = time_mfma_peak<double>(st, e0, e1, res, n_wg, 4096, &t64)) != SYMGW_OK) what = "FP64 MFMA peak";
      if (!r && (r = time_mfma_peak<float>(st, e0, e1, res, n_wg, 8192, &t32)) != SYMGW_OK) what = "FP32 MFMA peak";
      out[0] = std::max(out[0], t64);
      out[1] = std::max(out[1], t32);
    }
    if (!r && hipMemsetAsync(src, 0, (size_t)cbytes, st) != hipSuccess) r = SYMGW_ERR_HIP;
    for (int rep = 0; rep < 3 && !r; ++rep) {
      const int reps = 20, cwg = 64 * prop.multiProcessorCount;
      float ms = 0;
      if (launch_hbm_copy(src, dst, cbytes, cwg, st) != hipSuccess || hipEventRecord(e0, st) != hipSuccess)
        r = SYMGW_ERR_HIP;
      for (int i = 0; !r && i < reps; ++i)
        if (launch_hbm_copy(src, dst, cbytes, cwg, st) != hipSuccess) r = SYMGW_ERR_HIP;
      if (!r && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                 hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0))
        r = SYMGW_ERR_HIP;
      if (r) what = "HBM copy";
      else out[2] = std::max(out[2], 2.0 * (double)cbytes * reps / (ms * 1e-3) / 1e9);  // bytes read + written
    }
    if (st) (void)hipStreamSynchronize(st);
    if (res) dev_free(res);
    if (src) dev_free(src);
    if (dst) dev_free(dst);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (r) {
      (void)hipGetLastError();
      g_create_error = what + " failed";
    }
    return r;
  });
}

int symgw_create(const symgw_layout_desc *desc, int device, int precision, symgw_ctx **out) {
  return guarded(nullptr, [&]() -> int {
    if (!out) return SYMGW_ERR_ARG;
    *out = nullptr;
    if (!valid_precision(precision)) {
      g_create_error = "unknown precision mode";
      return SYMGW_ERR_ARG;
    }
    std::unique_ptr<symgw_ctx> c(new symgw_ctx());
    c->prec = prec_mode(precision);
    c->stream = prec_stream(precision);
    c->eb = precision_bytes(precision);
    c->shard_first = 0;
    std::string err;
    if (!build_layout(desc, c->L, err) ||
        !build_worklists(c->L, c->eb, c->W, err, c->prec == SYMGW_PREC_MIXED, c->stream)) {
      g_create_error = err;
      return SYMGW_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      g_create_error = "no HIP device";
      return SYMGW_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) {
      g_create_error = "device index out of range";
      return SYMGW_ERR_ARG;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      g_create_error = std::string("device is not gfx950: ") + prop.gcnArchName;
      return SYMGW_ERR_NODEVICE;
    }
    c->device = device;
    if (c->W.p0_lds > 160 * 1024 || *std::max_element(c->W.sig_lds, c->W.sig_lds + kNumSigDepths) > 160 * 1024) {
      g_create_error = "LDS plan exceeds 160 KiB";
      return SYMGW_ERR_ARG;
    }
    // the W / Dyson solvers keep one pivot per row in fixed-size LDS arrays
    if (c->L.max_ao_block > 512 || c->L.max_aux_block > 512) {
      g_create_error = "irrep blocks larger than 512 are not supported by the solvers";
      return SYMGW_ERR_ARG;
    }
    {
      std::lock_guard<std::mutex> lk(g_dev_mu);
      if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "stream creation failed";
        return SYMGW_ERR_HIP;
      }
      for (auto &e : c->ev) (void)hipEventCreate(&e);
    }
    const Layout &L = c->L;
    c->shard_end = L.ink;
    symgw_ctx *cp = c.get();
    int r = 0;
    const size_t eb = c->eb;
    r = r ? r : dalloc_bytes(cp, &c->V, (size_t)(c->stream ? c->W.batch_V_max : L.flat_V) * eb);
    r = r ? r : dalloc(cp, &c->G_ibz, (size_t)L.nts * L.ns * L.flat_ao);
    r = r ? r : dalloc_bytes(cp, &c->G_full, (size_t)L.nts * L.ns * L.flat_ao_full * eb);
    r = r ? r : dalloc(cp, &c->P, (size_t)L.nts * L.flat_aux);
    r = r ? r : dalloc_bytes(cp, &c->P_full, (size_t)L.nts * L.flat_aux_full * eb);
    r = r ? r : dalloc(cp, &c->Sig, (size_t)L.nts * L.ns * L.flat_ao);
    r = r ? r : dalloc(cp, &c->Pw, (size_t)std::max(1, L.nw_b) * L.flat_aux);
    r = r ? r : dalloc(cp, &c->SGw, (size_t)std::max(1, L.nw) * L.ns * L.flat_ao);
    r = r ? r : dalloc(cp, &c->F, (size_t)L.ns * L.flat_ao);
    r = r ? r : dalloc(cp, &c->S, (size_t)L.ns * L.flat_ao);
    r = r ? r : dalloc(cp, &c->orep, (size_t)L.flat_orep);
    r = r ? r : dalloc(cp, &c->auxrep, (size_t)L.flat_auxrep);
    r = r ? r : dalloc(cp, &c->Tnt, (size_t)L.nw * (L.nts - 2));
    r = r ? r : dalloc(cp, &c->Ttn, (size_t)L.nts * L.nw);
    r = r ? r : dalloc(cp, &c->Tnt_BF, (size_t)L.nw_b * (L.nts - 2));
    r = r ? r : dalloc(cp, &c->Ttn_FB, (size_t)L.nts * L.nw_b);
    r = r ? r : dalloc(cp, &c->iwn, (size_t)L.nw);
    r = r ? r : dalloc(cp, &c->w_scratch, (size_t)std::max(1, L.nw_b) * c->W.w_scr_group);
    r = r ? r : dalloc(cp, &c->d_scratch, (size_t)std::max(1, L.nw) * L.ns * c->W.d_scr_group);
    r = r ? r : dalloc(cp, &c->kw, (size_t)L.ink);
    r = r ? r : dalloc(cp, &c->nel_dev, 1);
    if (c->prec == SYMGW_PREC_MIXED) {
      r = r ? r : dalloc(cp, &c->V8, (size_t)c->W.v8_bytes);
      r = r ? r : dalloc(cp, &c->P8, (size_t)c->W.p8_bytes);
      r = r ? r : dalloc(cp, &c->VE, (size_t)c->W.ve_count);
      r = r ? r : dalloc(cp, &c->PE, (size_t)c->W.pe_count);
    }
    // every data buffer starts at zero (outside any graph: plain memsets are fine here)
    for (auto &b : c->owned)
      if (!r && hipMemsetAsync(b.p, 0, b.bytes, c->st) != hipSuccess) r = fail(cp, SYMGW_ERR_HIP, "initial memset failed");
    if (!r && hipStreamSynchronize(c->st) != hipSuccess) r = fail(cp, SYMGW_ERR_HIP, "initial memset failed");
    // work lists (the contraction tables come with the split tables, below)
    if (c->stream) {
      std::vector<GenBlk> g0, g1;
      batch_gen_blocks(L, c->W.p0_batch_pairs, c->W.p0_batch_pair_begin, g0, c->p0_gen_begin);
      batch_gen_blocks(L, c->W.sig_batch_pairs, c->W.sig_batch_pair_begin, g1, c->sig_gen_begin);
      r = r ? r : upload(cp, &c->d_p0_gen, g0);
      r = r ? r : upload(cp, &c->d_sig_gen, g1);
    }
    r = r ? r : upload_plan_tables(cp);
    r = r ? r : upload(cp, &c->d_g_rot, c->W.g_rot);
    r = r ? r : upload(cp, &c->d_p_rot, c->W.p_rot);
    r = r ? r : upload(cp, &c->d_w_blocks, c->W.w_blocks);
    r = r ? r : upload(cp, &c->d_d_blocks, c->W.d_blocks);
    // rotation source/destination slices, per full-BZ target
    {
      std::vector<int64_t> gs(L.nk), gsl(L.nk), gd(L.nk), gds(L.nk), ps(L.nk), psl(L.nk), pd(L.nk), pds(L.nk);
      for (int k = 0; k < L.nk; ++k) {
        int ick = L.irre_index[L.conj_index[k]];
        gs[k] = L.ns * L.flat_ao;  // G_ibz is ts_kij: one (t,s) row holds every iBZ k
        gsl[k] = L.kao_off_irre[ick];
        gd[k] = L.kao_off[k];
        gds[k] = L.kao_size[k];
        ps[k] = L.qaux_size_irre[ick];  // P is q_t_QP
        psl[k] = (int64_t)L.nts * L.qaux_off_irre[ick];
        pd[k] = L.qaux_off[k];
        pds[k] = L.qaux_size[k];
      }
      // G_ibz row stride is per (t, s): ts * flat_ao
      for (int k = 0; k < L.nk; ++k) gs[k] = L.flat_ao;
      r = r ? r : upload(cp, &c->d_g_src_stride, gs);
      r = r ? r : upload(cp, &c->d_g_src_slice, gsl);
      r = r ? r : upload(cp, &c->d_g_dst_slice, gd);
      r = r ? r : upload(cp, &c->d_g_dst_size, gds);
      r = r ? r : upload(cp, &c->d_p_src_stride, ps);
      r = r ? r : upload(cp, &c->d_p_src_slice, psl);
      r = r ? r : upload(cp, &c->d_p_dst_slice, pd);
      r = r ? r : upload(cp, &c->d_p_dst_size, pds);
    }
    {
      std::vector<int64_t> woff(L.ink), qsl(L.ink), qof(L.ink), doff(L.ink), ksl(L.ink), kof(L.ink);
      std::vector<TBatch> wf(L.ink), wb(L.ink), df(L.ink), db(L.ink);
      for (int i = 0; i < L.ink; ++i) {
        woff[i] = (int64_t)L.nw_b * L.qaux_off_irre[i];
        qsl[i] = L.qaux_size_irre[i];
        qof[i] = L.qaux_off_irre[i];
        doff[i] = (int64_t)L.nw * L.ns * L.kao_off_irre[i];
        ksl[i] = L.kao_size_irre[i];
        kof[i] = L.kao_off_irre[i];
        c->max_qslice = std::max<int>(c->max_qslice, (int)qsl[i]);
        c->max_kslice = std::max<int>(c->max_kslice, (int)ksl[i]);
        // W forward: P (q_t_QP) -> Pw ([q][n][qslice])
        wf[i] = TBatch{(int64_t)L.nts * L.qaux_off_irre[i], woff[i], qsl[i], qsl[i], 0, 0, (int32_t)qsl[i], 0};
        // W back: Pw -> P
        wb[i] = TBatch{woff[i], (int64_t)L.nts * L.qaux_off_irre[i], qsl[i], qsl[i], 0, 0, (int32_t)qsl[i], 0};
        // Dyson forward: Sig (k_ts_ij) -> SGw ([k][w][s][kslice]); (s, x) contiguous
        df[i] = TBatch{(int64_t)L.nts * L.ns * L.kao_off_irre[i], doff[i], L.ns * ksl[i], L.ns * ksl[i], 0, 0,
                       (int32_t)(L.ns * ksl[i]), 0};
        // Dyson back: SGw -> G_ibz (ts_kij): columns split per spin
        db[i] = TBatch{doff[i], L.kao_off_irre[i], L.ns * ksl[i], L.ns * L.flat_ao, ksl[i], L.flat_ao,
                       (int32_t)ksl[i], 0};
      }
      r = r ? r : upload(cp, &c->d_w_off, woff);
      r = r ? r : upload(cp, &c->d_qslice_irre, qsl);
      r = r ? r : upload(cp, &c->d_qoff_irre, qof);
      r = r ? r : upload(cp, &c->d_d_off, doff);
      r = r ? r : upload(cp, &c->d_kslice_irre, ksl);
      r = r ? r : upload(cp, &c->d_kao_off_irre, kof);
      r = r ? r : upload(cp, &c->d_tb_wf, wf);
      r = r ? r : upload(cp, &c->d_tb_wb, wb);
      r = r ? r : upload(cp, &c->d_tb_df, df);
      r = r ? r : upload(cp, &c->d_tb_db, db);
    }
    if (r) {
      g_create_error = c->err;
      release_ctx(c.get());
      return r;
    }
    *out = c.release();
    return SYMGW_OK;
  });
}

int symgw_destroy(symgw_ctx *ctx) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->st);
    release_ctx(ctx);
    delete ctx;
    return SYMGW_OK;
  });
}

int symgw_layout_info_ctx(const symgw_ctx *ctx, symgw_layout_info *info) {
  return guarded(const_cast<symgw_ctx *>(ctx), [&]() -> int {
    if (!ctx || !info) return SYMGW_ERR_ARG;
    const Layout &L = ctx->L;
    info->flat_ao_size = L.flat_ao;
    info->flat_ao_size_full = L.flat_ao_full;
    info->flat_aux_size = L.flat_aux;
    info->flat_aux_size_full = L.flat_aux_full;
    info->flat_V_size = L.flat_V;
    info->flat_orep_size = L.flat_orep;
    info->flat_auxrep_size = L.flat_auxrep;
    info->n_P0_blocks = ctx->W.n_p0_contractions;
    info->n_Sigma_blocks = ctx->W.n_sig_contractions;
    info->flop_P0 = ctx->W.flop_p0;
    info->flop_Sigma = ctx->W.flop_sig;
    info->device_bytes = footprint(L, ctx->W, ctx->eb);
    return SYMGW_OK;
  });
}

int64_t symgw_kpair_offset(const symgw_ctx *ctx, int k1, int k2) {
  if (!ctx || k1 < 0 || k2 < 0 || k1 >= ctx->L.nk || k2 >= ctx->L.nk) return -1;
  size_t p = (size_t)k1 * ctx->L.nk + k2;
  return ctx->L.V_irreps[p].empty() ? -1 : ctx->L.kpair_off[p];
}

int symgw_set_V(symgw_ctx *ctx, const double *V_flat, int64_t n_complex) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !V_flat) return SYMGW_ERR_ARG;
    if (ctx->stream) return fail(ctx, SYMGW_ERR_STATE, "a context with a streamed V takes symgw_generate_V only");
    if (n_complex != ctx->L.flat_V) return fail(ctx, SYMGW_ERR_ARG, "V length does not match the layout");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (ctx->eb == 16) {
      HIPCHK(ctx, copy_sync(ctx, ctx->V, V_flat, (size_t)n_complex * 16, hipMemcpyHostToDevice));
    } else {
      // staged conversion in 64 MiB chunks
      const int64_t chunk = 4 << 20;
      cplx<double> *stage = nullptr;
      HIPCHK(ctx, ctx_malloc(ctx, (void **)&stage, (size_t)chunk * 16));
      for (int64_t o = 0; o < n_complex; o += chunk) {
        int64_t n = std::min(chunk, n_complex - o);
        hipError_t e = copy_sync(ctx, stage, V_flat + 2 * o, (size_t)n * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = launch_convert<float>(stage, static_cast<cplx<float> *>(ctx->V) + o, n, ctx->st);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
        if (e != hipSuccess) {
          ctx_free(ctx, stage);
          return fail(ctx, SYMGW_ERR_HIP, std::string("V upload: ") + hipGetErrorString(e));
        }
      }
      ctx_free(ctx, stage);
    }
    int r = split_V8(ctx);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    ctx->have_V = true;
    return SYMGW_OK;
  });
}

static void gen_blocks(const Layout &L, std::vector<GenBlk> &out, int only_pair) {
  for (size_t p = 0; p < L.V_irreps.size(); ++p) {
    if (only_pair >= 0 && (int)p != only_pair) continue;
    int k1 = (int)(p / L.nk), k2 = (int)(p % L.nk);
    for (size_t b = 0; b < L.V_irreps[p].size(); ++b) {
      auto &t = L.V_irreps[p][b];
      GenBlk g;
      g.v_off = L.kpair_off[p] + L.V_offsets[p][b];
      g.local_off = L.V_offsets[p][b];
      g.q = L.V_qsize[p][b];
      g.k1 = L.ao_sizes[k1][t[1]];
      g.k2 = L.ao_sizes[k2][t[2]];
      g.a_base = L.ao_block_offsets[k1][t[1]];
      g.b_base = L.ao_block_offsets[k2][t[2]];
      g.pair = (int32_t)p;
      out.push_back(g);
    }
  }
}

int symgw_generate_V(symgw_ctx *ctx, uint64_t seed, double scale, const double *decay, int n_decay) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !decay || n_decay <= 0) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (ctx->stream) {
      // streamed V: keep the generator; every P0 / Sigma batch regenerates its pairs
      if (ctx->d_decay && ctx->n_decay != n_decay) return fail(ctx, SYMGW_ERR_ARG, "decay table length changed");
      if (!ctx->d_decay) {
        int r = dalloc(ctx, &ctx->d_decay, (size_t)n_decay);
        if (r) return r;
      }
      HIPCHK(ctx, copy_sync(ctx, ctx->d_decay, decay, (size_t)n_decay * sizeof(double), hipMemcpyHostToDevice));
      ctx->n_decay = n_decay;
      ctx->gen_seed = seed;
      ctx->gen_scale = scale;
      ctx->have_V = true;
      return SYMGW_OK;
    }
    std::vector<GenBlk> gb;
    gen_blocks(ctx->L, gb, -1);
    GenBlk *dgb = nullptr;
    double *ddec = nullptr;
    HIPCHK(ctx, ctx_malloc(ctx, (void **)&dgb, std::max<size_t>(1, gb.size()) * sizeof(GenBlk)));
    hipError_t e0 = ctx_malloc(ctx, (void **)&ddec, (size_t)n_decay * sizeof(double));
    if (e0 != hipSuccess) {
      ctx_free(ctx, dgb);
      return fail(ctx, SYMGW_ERR_NOMEM, "V generation: allocation failed");
    }
    hipError_t e = copy_sync(ctx, dgb, gb.data(), gb.size() * sizeof(GenBlk), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = copy_sync(ctx, ddec, decay, (size_t)n_decay * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      e = ctx->eb == 16
              ? launch_genv<double>(dgb, (int)gb.size(), seed, scale, ddec, n_decay, static_cast<cplx<double> *>(ctx->V),
                                    ctx->st)
              : launch_genv<float>(dgb, (int)gb.size(), seed, scale, ddec, n_decay, static_cast<cplx<float> *>(ctx->V),
                                   ctx->st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
    ctx_free(ctx, dgb);
    ctx_free(ctx, ddec);
    if (e != hipSuccess) return fail(ctx, SYMGW_ERR_HIP, std::string("V generation: ") + hipGetErrorString(e));
    int r = split_V8(ctx);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    ctx->have_V = true;
    return SYMGW_OK;
  });
}

int symgw_generate_V_pair_host(const symgw_layout_desc *desc, uint64_t seed, double scale, const double *decay,
                               int n_decay, int k1, int k2, double *out) {
  return guarded(nullptr, [&]() -> int {
    if (!desc || !decay || n_decay <= 0 || !out) return SYMGW_ERR_ARG;
    Layout L;
    std::string err;
    if (!build_layout(desc, L, err)) {
      g_create_error = err;
      return SYMGW_ERR_ARG;
    }
    if (k1 < 0 || k2 < 0 || k1 >= L.nk || k2 >= L.nk) return SYMGW_ERR_ARG;
    std::vector<GenBlk> gb;
    gen_blocks(L, gb, k1 * L.nk + k2);
    for (const GenBlk &B : gb) {
      const int64_t K = (int64_t)B.k1 * B.k2, n = (int64_t)B.q * K;
      for (int64_t e = 0; e < n; ++e) {
        const int64_t rem = e % K;
        const int a = (int)(rem / B.k2), m = (int)(rem % B.k2);
        int d = (B.a_base + a) - (B.b_base + m);
        d = d < 0 ? -d : d;
        d = d < n_decay ? d : n_decay - 1;
        const double f = (scale * decay[d]) * 0.70710678118654752440;
        double re, im;
        gen_v_elem(seed, (uint32_t)B.pair, (uint64_t)(B.local_off + e), f, re, im);
        out[2 * (B.local_off + e)] = re;
        out[2 * (B.local_off + e) + 1] = im;
      }
    }
    return SYMGW_OK;
  });
}

int symgw_generate_V_pairs_host(const symgw_layout_desc *desc, uint64_t seed, double scale, const double *decay,
                                int n_decay, int npairs, const int32_t *pairs, const int64_t *out_off, double *out) {
  return guarded(nullptr, [&]() -> int {
    if (!desc || !decay || n_decay <= 0 || npairs < 0 || (npairs > 0 && (!pairs || !out_off || !out)))
      return SYMGW_ERR_ARG;
    Layout L;
    std::string err;
    if (!build_layout(desc, L, err)) {
      g_create_error = err;
      return SYMGW_ERR_ARG;
    }
    for (int i = 0; i < npairs; ++i) {
      const int k1 = pairs[2 * i], k2 = pairs[2 * i + 1];
      if (k1 < 0 || k2 < 0 || k1 >= L.nk || k2 >= L.nk || L.V_irreps[(size_t)k1 * L.nk + k2].empty()) {
        g_create_error = "generate_V_pairs_host: pair not stored";
        return SYMGW_ERR_ARG;
      }
    }
    // pairs are independent: spread them over the host threads
    const int nth = std::max(1, std::min<int>(npairs, (int)std::thread::hardware_concurrency()));
    std::atomic<int> next(0);
    auto work = [&]() {
      std::vector<GenBlk> gb;
      for (int i = next++; i < npairs; i = next++) {
        gb.clear();
        gen_blocks(L, gb, pairs[2 * i] * L.nk + pairs[2 * i + 1]);
        double *o = out + 2 * out_off[i];
        for (const GenBlk &B : gb) {
          const int64_t K = (int64_t)B.k1 * B.k2, n = (int64_t)B.q * K;
          for (int64_t e = 0; e < n; ++e) {
            const int64_t rem = e % K;
            const int a = (int)(rem / B.k2), m = (int)(rem % B.k2);
            int d = (B.a_base + a) - (B.b_base + m);
            d = d < 0 ? -d : d;
            d = d < n_decay ? d : n_decay - 1;
            const double f = (scale * decay[d]) * 0.70710678118654752440;
            double re, im;
            gen_v_elem(seed, (uint32_t)B.pair, (uint64_t)(B.local_off + e), f, re, im);
            o[2 * (B.local_off + e)] = re;
            o[2 * (B.local_off + e) + 1] = im;
          }
        }
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    return SYMGW_OK;
  });
}

int symgw_set_reps(symgw_ctx *ctx, const double *orep_flat, const double *auxrep_flat) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (ctx->L.flat_orep > 0) {
      if (!orep_flat) return fail(ctx, SYMGW_ERR_ARG, "orep missing");
      HIPCHK(ctx, copy_sync(ctx, ctx->orep, orep_flat, (size_t)ctx->L.flat_orep * 16, hipMemcpyHostToDevice));
    }
    if (ctx->L.flat_auxrep > 0) {
      if (!auxrep_flat) return fail(ctx, SYMGW_ERR_ARG, "auxrep missing");
      HIPCHK(ctx, copy_sync(ctx, ctx->auxrep, auxrep_flat, (size_t)ctx->L.flat_auxrep * 16, hipMemcpyHostToDevice));
    }
    ctx->have_reps = true;
    ctx->p_rotated = false;
    return SYMGW_OK;
  });
}

int symgw_upload_G(symgw_ctx *ctx, const double *G_ts_kij) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !G_ts_kij) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Layout &L = ctx->L;
    HIPCHK(ctx, copy_sync(ctx, ctx->G_ibz, G_ts_kij, (size_t)L.nts * L.ns * L.flat_ao * 16, hipMemcpyHostToDevice));
    ctx->have_G = true;
    int r = rotate_G(ctx);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return SYMGW_OK;
  });
}

int symgw_download_G(symgw_ctx *ctx, double *G_ts_kij) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !G_ts_kij) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    const Layout &L = ctx->L;
    HIPCHK(ctx, copy_sync(ctx, G_ts_kij, ctx->G_ibz, (size_t)L.nts * L.ns * L.flat_ao * 16, hipMemcpyDeviceToHost));
    return SYMGW_OK;
  });
}

int symgw_solve_P0(symgw_ctx *ctx) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    ev_record(ctx, 0);
    int r = phase_P0(ctx);
    if (r) return r;
    ev_record(ctx, 1);
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    if (ctx->timing) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
      ctx->last_ms[0] = ms;
    }
    ctx->have_P = false;
    return SYMGW_OK;
  });
}

int symgw_copy_P0_to_host(const symgw_ctx *ctx, double *P0_q_t_QP) {
  return guarded(const_cast<symgw_ctx *>(ctx), [&]() -> int {
    if (!ctx || !P0_q_t_QP) return SYMGW_ERR_ARG;
    symgw_ctx *c = const_cast<symgw_ctx *>(ctx);
    HIPCHK(c, hipSetDevice(ctx->device));
    HIPCHK(c, hipStreamSynchronize(ctx->st));
    const size_t n = (size_t)ctx->L.nts * ctx->L.flat_aux;
    std::vector<double> tmp(2 * n);
    HIPCHK(c, copy_sync(c, tmp.data(), ctx->P, n * 16, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < 2 * n; ++i) P0_q_t_QP[i] += tmp[i];
    return SYMGW_OK;
  });
}

int symgw_upload_P(symgw_ctx *ctx, const double *P_q_t_QP) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !P_q_t_QP) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, copy_sync(ctx, ctx->P, P_q_t_QP, (size_t)ctx->L.nts * ctx->L.flat_aux * 16, hipMemcpyHostToDevice));
    ctx->have_P = true;
    ctx->p_rotated = false;
    return SYMGW_OK;
  });
}

int symgw_set_transforms(symgw_ctx *ctx, const double *Tnt, const double *Ttn, const double *Tnt_BF,
                         const double *Ttn_FB, const double *iwn_fermi) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !Tnt || !Ttn || !Tnt_BF || !Ttn_FB || !iwn_fermi) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Layout &L = ctx->L;
    HIPCHK(ctx, copy_sync(ctx, ctx->Tnt, Tnt, (size_t)L.nw * (L.nts - 2) * 16, hipMemcpyHostToDevice));
    HIPCHK(ctx, copy_sync(ctx, ctx->Ttn, Ttn, (size_t)L.nts * L.nw * 16, hipMemcpyHostToDevice));
    HIPCHK(ctx, copy_sync(ctx, ctx->Tnt_BF, Tnt_BF, (size_t)L.nw_b * (L.nts - 2) * 16, hipMemcpyHostToDevice));
    HIPCHK(ctx, copy_sync(ctx, ctx->Ttn_FB, Ttn_FB, (size_t)L.nts * L.nw_b * 16, hipMemcpyHostToDevice));
    HIPCHK(ctx, copy_sync(ctx, ctx->iwn, iwn_fermi, (size_t)L.nw * 16, hipMemcpyHostToDevice));
    ctx->have_T = true;
    return SYMGW_OK;
  });
}

int symgw_solve_W(symgw_ctx *ctx) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int r = phase_W(ctx);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return SYMGW_OK;
  });
}

int symgw_solve_W_noalloc(symgw_ctx *ctx) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices)
      return fail(ctx, SYMGW_ERR_ARG, "solve_W_noalloc: device index above the supported range");
    const int64_t cap = (int64_t)(kNoAllocScratchBytes / sizeof(cplx<double>));
    const int64_t group = ctx->W.w_scr_group;  // complex elements of one bosonic frequency's solves
    if (group > cap)
      return fail(ctx, SYMGW_ERR_ARG,
                  "solve_W_noalloc: the solves of one bosonic frequency need " + std::to_string(group * 16) +
                      " bytes of scratch, above the 32 MiB of the allocation-free path");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    void *&ptr = g_noalloc_scratch[ctx->device];
    if (!ptr) HIPCHK(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    int r = phase_W_with(ctx, static_cast<cplx<double> *>(ptr), std::max<int64_t>(1, cap / std::max<int64_t>(1, group)));
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return SYMGW_OK;
  });
}

int symgw_solve_Sigma(symgw_ctx *ctx) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    ev_record(ctx, 2);
    int r = phase_Sigma(ctx);
    if (r) return r;
    ev_record(ctx, 3);
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    if (ctx->timing) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
      ctx->last_ms[2] = ms;
    }
    return SYMGW_OK;
  });
}

int symgw_copy_Sigma_to_host(const symgw_ctx *ctx, double *Sigma_ts_kij) {
  return guarded(const_cast<symgw_ctx *>(ctx), [&]() -> int {
    if (!ctx || !Sigma_ts_kij) return SYMGW_ERR_ARG;
    symgw_ctx *c = const_cast<symgw_ctx *>(ctx);
    HIPCHK(c, hipSetDevice(ctx->device));
    HIPCHK(c, hipStreamSynchronize(ctx->st));
    const Layout &L = ctx->L;
    const size_t n = (size_t)L.nts * L.ns * L.flat_ao;
    std::vector<double> tmp(2 * n);
    HIPCHK(c, copy_sync(c, tmp.data(), ctx->Sig, n * 16, hipMemcpyDeviceToHost));
    // k_ts_ij -> ts_kij with accumulation (cugw_routines_block.cu:257-273)
    const int nts_ns = L.nts * L.ns;
    for (int ik = 0; ik < L.ink; ++ik) {
      const int64_t sz = L.kao_size_irre[ik], off = L.kao_off_irre[ik];
      for (int ts = 0; ts < nts_ns; ++ts) {
        const double *src = tmp.data() + 2 * (off * nts_ns + (int64_t)ts * sz);
        double *dst = Sigma_ts_kij + 2 * ((int64_t)ts * L.flat_ao + off);
        for (int64_t x = 0; x < 2 * sz; ++x) dst[x] += src[x];
      }
    }
    return SYMGW_OK;
  });
}

int symgw_upload_Sigma(symgw_ctx *ctx, const double *Sigma_ts_kij) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !Sigma_ts_kij) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Layout &L = ctx->L;
    const size_t n = (size_t)L.nts * L.ns * L.flat_ao;
    std::vector<double> tmp(2 * n);
    const int nts_ns = L.nts * L.ns;
    for (int ik = 0; ik < L.ink; ++ik) {
      const int64_t sz = L.kao_size_irre[ik], off = L.kao_off_irre[ik];
      for (int ts = 0; ts < nts_ns; ++ts) {
        const double *src = Sigma_ts_kij + 2 * ((int64_t)ts * L.flat_ao + off);
        double *dst = tmp.data() + 2 * (off * nts_ns + (int64_t)ts * sz);
        std::memcpy(dst, src, (size_t)sz * 16);
      }
    }
    HIPCHK(ctx, copy_sync(ctx, ctx->Sig, tmp.data(), n * 16, hipMemcpyHostToDevice));
    ctx->have_Sig = true;
    return SYMGW_OK;
  });
}

int symgw_set_fock(symgw_ctx *ctx, const double *F, const double *S) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !F || !S) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->L.ns * ctx->L.flat_ao;
    HIPCHK(ctx, copy_sync(ctx, ctx->F, F, n * 16, hipMemcpyHostToDevice));
    HIPCHK(ctx, copy_sync(ctx, ctx->S, S, n * 16, hipMemcpyHostToDevice));
    ctx->have_F = true;
    return SYMGW_OK;
  });
}

int symgw_dyson(symgw_ctx *ctx, double mu, double sigma_scale) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int r = phase_Dyson(ctx, mu, sigma_scale);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return SYMGW_OK;
  });
}

int symgw_density(symgw_ctx *ctx, const double *kweights, double *nel) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !kweights || !nel) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (!ctx->have_G || !ctx->have_F) return fail(ctx, SYMGW_ERR_STATE, "G or S missing");
    const Layout &L = ctx->L;
    HIPCHK(ctx, hipMemcpyAsync(ctx->kw, kweights, (size_t)L.ink * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    const cplx<double> *Gbeta = ctx->G_ibz + (int64_t)(L.nts - 1) * L.ns * L.flat_ao;
    HIPCHK(ctx, launch_density(ctx->d_d_blocks, (int)ctx->W.d_blocks.size(), L.ns, Gbeta, ctx->S, ctx->d_kao_off_irre,
                               L.flat_ao, ctx->kw, ctx->nel_dev, ctx->st));
    double v = 0;
    HIPCHK(ctx, hipMemcpyAsync(&v, ctx->nel_dev, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    // dmr = -G(beta) * (2 / ns) (sc_loop_gf_block.cpp:18-19), per k weighted by w_k / nk
    *nel = -(2.0 / L.ns) * v / L.nk;
    return SYMGW_OK;
  });
}

int symgw_mu_search(symgw_ctx *ctx, double nel_target, const double *kweights, double sigma_scale, double mu_lo,
                    double mu_hi, double tol, double *mu_out, double *nel_out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !kweights || !mu_out || mu_hi <= mu_lo || tol <= 0) return SYMGW_ERR_ARG;
    double lo = mu_lo, hi = mu_hi, mu = 0.5 * (lo + hi), n = 0, nlo = 0, nhi = 0;
    int r;
    auto count = [&](double m, double &nn) {
      int rr = symgw_dyson(ctx, m, sigma_scale);
      return rr ? rr : symgw_density(ctx, kweights, &nn);
    };
    // widen [lo, hi] (doubling its width on the short side) until it brackets the target
    if ((r = count(lo, nlo)) || (r = count(hi, nhi))) return r;
    for (int ex = 0; ex < 8 && (nlo > nel_target || nhi < nel_target); ++ex) {
      const double w = hi - lo;
      if (nlo > nel_target && (r = count(lo -= w, nlo))) return r;
      if (nhi < nel_target && (r = count(hi += w, nhi))) return r;
    }
    if (nlo > nel_target || nhi < nel_target)
      return fail(ctx, SYMGW_ERR_ARG, "mu_search: the electron count is not bracketed");
    for (int it = 0; it < 200; ++it) {
      mu = 0.5 * (lo + hi);
      if ((r = symgw_dyson(ctx, mu, sigma_scale))) return r;
      if ((r = symgw_density(ctx, kweights, &n))) return r;
      // N(mu) increases with mu
      if (n < nel_target) lo = mu; else hi = mu;
      if (std::fabs(n - nel_target) < tol || hi - lo < 1e-14) break;
    }
    *mu_out = mu;
    if (nel_out) *nel_out = n;
    return SYMGW_OK;
  });
}

int symgw_gw_iteration(symgw_ctx *ctx, double mu, int use_graph) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (!ctx->have_V || !ctx->have_G || !ctx->have_T || !ctx->have_F)
      return fail(ctx, SYMGW_ERR_STATE, "V, G, transforms and Fock must be set first");
    ctx->have_P = true;   // produced inside the sequence
    ctx->have_Sig = true;
    if (!use_graph) {
      int r = iteration_body(ctx, mu);
      if (r) return r;
    } else {
      std::unique_lock<std::mutex> lk(g_dev_mu);
      if (ctx->graph_exec[0] && ctx->graph_mu != mu) drop_graphs(ctx);
      if (!ctx->graph_exec[0]) {
        const bool tm = ctx->timing;
        ctx->timing = false;
        for (int ph = 0; ph < 4; ++ph) {
          hipGraph_t g = nullptr;
          hipError_t e = hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal);
          if (e != hipSuccess) {
            ctx->timing = tm;
            drop_graphs(ctx);
            return fail(ctx, SYMGW_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
          }
          int r = run_phase(ctx, ph, mu);
          e = hipStreamEndCapture(ctx->st, &g);
          if (!r && e == hipSuccess) e = hipGraphInstantiate(&ctx->graph_exec[ph], g, nullptr, nullptr, 0);
          if (g) (void)hipGraphDestroy(g);
          if (r || e != hipSuccess) {
            ctx->timing = tm;
            drop_graphs(ctx);
            return r ? r : fail(ctx, SYMGW_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
          }
        }
        ctx->timing = tm;
        ctx->graph_mu = mu;
      }
      lk.unlock();
      ev_record(ctx, 0);
      for (int ph = 0; ph < 4; ++ph) {
        HIPCHK(ctx, hipGraphLaunch(ctx->graph_exec[ph], ctx->st));
        ev_record(ctx, ph + 1);
      }
    }
    return SYMGW_OK;
  });
}

int symgw_set_shard(symgw_ctx *ctx, int first, int count) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    if (first < 0 || count < 1 || first + count > ctx->L.ink)
      return fail(ctx, SYMGW_ERR_ARG, "shard outside the irreducible points");
    ctx->shard_first = first;
    ctx->shard_end = first + count;
    std::lock_guard<std::mutex> lk(g_dev_mu);
    drop_graphs(ctx);  // captured sweeps cover the previous shard
    return SYMGW_OK;
  });
}

int symgw_plan_defaults(symgw_plan *plan) {
  return guarded(nullptr, [&]() -> int {
    if (!plan) return SYMGW_ERR_ARG;
    const Plan d;
    plan->sig_chunk_cost = d.sig_chunk_cost;
    plan->sig_rows = d.sig_rows;
    plan->sig_max_times = d.sig_max_times;
    plan->p0_chunk_cost = d.p0_chunk_cost;
    plan->p0_band_cap = d.p0_band_cap;
    plan->reserved = 0;
    return SYMGW_OK;
  });
}

int symgw_get_plan(const symgw_ctx *ctx, symgw_plan *plan) {
  return guarded(nullptr, [&]() -> int {
    if (!ctx || !plan) return SYMGW_ERR_ARG;
    plan->sig_chunk_cost = ctx->plan.sig_chunk_cost;
    plan->sig_rows = ctx->plan.sig_rows;
    plan->sig_max_times = ctx->plan.sig_max_times;
    plan->p0_chunk_cost = ctx->plan.p0_chunk_cost;
    plan->p0_band_cap = ctx->plan.p0_band_cap;
    plan->reserved = 0;
    return SYMGW_OK;
  });
}

int symgw_replan(symgw_ctx *ctx, const symgw_plan *plan) {
  return guarded(ctx, [&]() -> int {
    if (!ctx || !plan) return SYMGW_ERR_ARG;
    const Plan p = plan_of(*plan);
    if (!valid_plan(p)) return fail(ctx, SYMGW_ERR_ARG, "plan parameters out of range");
    WorkLists W2;
    std::string err;
    if (!build_worklists(ctx->L, ctx->eb, W2, err, ctx->prec == SYMGW_PREC_MIXED, ctx->stream, p))
      return fail(ctx, SYMGW_ERR_ARG, err);
    const WorkLists &W = ctx->W;
    // the data buffers sized from the work lists must stay valid
    if (W2.batch_V_max != W.batch_V_max || W2.v8_bytes != W.v8_bytes || W2.p8_bytes != W.p8_bytes ||
        W2.ve_count != W.ve_count || W2.pe_count != W.pe_count || W2.w_scr_group != W.w_scr_group ||
        W2.d_scr_group != W.d_scr_group || W2.p0_group_first != W.p0_group_first ||
        W2.sig_group_first != W.sig_group_first || W2.v8_batch_begin != W.v8_batch_begin)
      return fail(ctx, SYMGW_ERR_ARG, "the plan changes buffer sizes of the context");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    {
      std::lock_guard<std::mutex> lk(g_dev_mu);
      drop_graphs(ctx);  // captured sweeps read the old tables
    }
    ctx->W = std::move(W2);
    ctx->plan = p;
    int r = upload_plan_tables(ctx);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return SYMGW_OK;
  });
}

int symgw_sync(symgw_ctx *ctx) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    if (ctx->timing) {
      for (int i = 0; i < 5; ++i) {
        float ms = 0;
        const int a = i < 4 ? i : 0, b = i < 4 ? i + 1 : 4;
        if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) == hipSuccess) ctx->last_ms[i] = ms;
        else (void)hipGetLastError();
      }
    }
    return SYMGW_OK;
  });
}

int symgw_last_times(const symgw_ctx *ctx, double *ms5) {
  return guarded(const_cast<symgw_ctx *>(ctx), [&]() -> int {
    if (!ctx || !ms5) return SYMGW_ERR_ARG;
    for (int i = 0; i < 5; ++i) ms5[i] = ctx->last_ms[i];
    return SYMGW_OK;
  });
}

int symgw_set_timing(symgw_ctx *ctx, int on) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return SYMGW_ERR_ARG;
    ctx->timing = on != 0;
    return SYMGW_OK;
  });
}

}  // extern "C"
