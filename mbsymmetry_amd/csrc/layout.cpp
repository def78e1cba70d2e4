// Derived block tables and device work lists.
//
// Table derivation follows symmetry_utils_t (SymmetryMBPT/src/common/
// symmetry_utils.cpp:96-191: cumulative b*b flat offsets per k, iBZ slices in
// irre_list order) and the kpair slice tables of df_integral_block_t
// (SymmetryMBPT/include/common/df_integral_block_t.h:84-106).
#include "layout.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <map>
#include <sstream>

namespace symgw {

namespace {
std::string fmt(const char *what, long a, long b = -1) {
  std::ostringstream o;
  o << what << " (" << a;
  if (b >= 0) o << ", " << b;
  o << ")";
  return o.str();
}
}  // namespace

bool build_layout(const symgw_layout_desc *d, Layout &L, std::string &err) {
  if (!d) { err = "null layout"; return false; }
  L = Layout();
  L.nk = d->nk; L.ink = d->ink; L.ns = d->ns; L.nts = d->nts; L.nw = d->nw; L.nw_b = d->nw_b;
  const int nk = L.nk, ink = L.ink;
  if (nk <= 0 || ink <= 0 || ink > nk) { err = fmt("bad nk/ink", nk, ink); return false; }
  if (L.ns != 1 && L.ns != 2) { err = fmt("ns must be 1 or 2", L.ns); return false; }
  // odd nts is rejected by the reference (cugw_solver_block.h:39-40)
  if (L.nts < 4 || (L.nts % 2) != 0) { err = fmt("nts must be even and >= 4", L.nts); return false; }
  if (L.nw < 0 || L.nw_b < 0) { err = "negative nw"; return false; }
  if (!d->kao_nblocks || !d->ao_sizes || !d->qaux_nblocks || !d->aux_sizes || !d->irre_list ||
      !d->index || !d->irre_index || !d->conj_index || !d->kpq || !d->kmq || !d->kpair_nblocks) {
    err = "missing layout array"; return false;
  }
  L.ao_sizes.resize(nk); L.aux_sizes.resize(nk);
  L.ao_offsets.resize(nk); L.aux_offsets.resize(nk);
  L.ao_block_offsets.resize(nk); L.aux_block_offsets.resize(nk);
  L.kao_off.resize(nk); L.kao_size.resize(nk); L.qaux_off.resize(nk); L.qaux_size.resize(nk);
  int64_t pa = 0, px = 0, ka = 0, kx = 0;
  for (int k = 0; k < nk; ++k) {
    int nb = d->kao_nblocks[k], nx = d->qaux_nblocks[k];
    if (nb <= 0 || nx <= 0) { err = fmt("empty block list at k", k); return false; }
    int64_t off = 0; int boff = 0;
    for (int b = 0; b < nb; ++b) {
      int s = d->ao_sizes[pa + b];
      if (s <= 0) { err = fmt("non-positive ao block size at k", k, b); return false; }
      L.ao_sizes[k].push_back(s); L.ao_offsets[k].push_back(off); L.ao_block_offsets[k].push_back(boff);
      off += (int64_t)s * s; boff += s;
      L.max_ao_block = std::max(L.max_ao_block, s);
    }
    pa += nb;
    L.kao_off[k] = ka; L.kao_size[k] = off; ka += off;
    off = 0; boff = 0;
    for (int b = 0; b < nx; ++b) {
      int s = d->aux_sizes[px + b];
      if (s <= 0) { err = fmt("non-positive aux block size at q", k, b); return false; }
      L.aux_sizes[k].push_back(s); L.aux_offsets[k].push_back(off); L.aux_block_offsets[k].push_back(boff);
      off += (int64_t)s * s; boff += s;
      L.max_aux_block = std::max(L.max_aux_block, s);
    }
    px += nx;
    L.qaux_off[k] = kx; L.qaux_size[k] = off; kx += off;
  }
  L.flat_ao_full = ka; L.flat_aux_full = kx;

  L.irre_list.assign(d->irre_list, d->irre_list + ink);
  L.index.assign(d->index, d->index + nk);
  L.irre_index.assign(d->irre_index, d->irre_index + nk);
  L.conj_index.assign(d->conj_index, d->conj_index + nk);
  L.kpq.assign(d->kpq, d->kpq + (size_t)nk * nk);
  L.kmq.assign(d->kmq, d->kmq + (size_t)nk * nk);
  for (int i = 0; i < ink; ++i)
    if (L.irre_list[i] < 0 || L.irre_list[i] >= nk) { err = fmt("irre_list out of range", i); return false; }
  for (int k = 0; k < nk; ++k) {
    if (L.index[k] < 0 || L.index[k] >= nk || L.irre_index[k] < 0 || L.irre_index[k] >= ink ||
        L.conj_index[k] < 0 || L.conj_index[k] >= nk) {
      err = fmt("symmetry map out of range at k", k); return false;
    }
    // time-reversal partners share their block structure (symmetry_utils.cpp:122-136)
    int ck = L.conj_index[k];
    if (L.ao_sizes[k] != L.ao_sizes[ck] || L.aux_sizes[k] != L.aux_sizes[ck]) {
      err = fmt("k and its conjugate partner differ in blocks", k, ck); return false;
    }
    for (int q = 0; q < nk; ++q) {
      int a = L.kpq[(size_t)k * nk + q], b = L.kmq[(size_t)k * nk + q];
      if (a < 0 || a >= nk || b < 0 || b >= nk) { err = fmt("kpq/kmq out of range", k, q); return false; }
    }
  }
  // iBZ slices in irre_list order (symmetry_utils.cpp:174-191)
  L.kao_off_irre.resize(ink); L.kao_size_irre.resize(ink);
  L.qaux_off_irre.resize(ink); L.qaux_size_irre.resize(ink);
  int64_t oa = 0, ox = 0;
  for (int i = 0; i < ink; ++i) {
    int k = L.irre_list[i];
    L.kao_off_irre[i] = oa; L.kao_size_irre[i] = L.kao_size[k]; oa += L.kao_size[k];
    L.qaux_off_irre[i] = ox; L.qaux_size_irre[i] = L.qaux_size[k]; ox += L.qaux_size[k];
  }
  L.flat_ao = oa; L.flat_aux = ox;

  // rotation blocks
  L.orep_irreps.resize(nk); L.auxrep_irreps.resize(nk);
  L.orep_offsets.resize(nk); L.auxrep_offsets.resize(nk);
  L.korep_off.resize(nk); L.korep_size.resize(nk); L.kauxrep_off.resize(nk); L.kauxrep_size.resize(nk);
  int64_t po = 0, pp = 0, ro = 0, rx = 0;
  for (int k = 0; k < nk; ++k) {
    int nbo = d->orep_nblocks ? d->orep_nblocks[k] : 0;
    int nbx = d->auxrep_nblocks ? d->auxrep_nblocks[k] : 0;
    int ik = L.irre_index[k];
    int kr = L.irre_list[ik];
    int64_t off = 0;
    for (int b = 0; b < nbo; ++b) {
      int a1 = d->orep_irreps[2 * (po + b)], a2 = d->orep_irreps[2 * (po + b) + 1];
      if (a1 < 0 || a1 >= (int)L.ao_sizes[k].size() || a2 < 0 || a2 >= (int)L.ao_sizes[kr].size()) {
        err = fmt("orep block index out of range at k", k, b); return false;
      }
      L.orep_irreps[k].push_back({a1, a2}); L.orep_offsets[k].push_back(off);
      off += (int64_t)L.ao_sizes[k][a1] * L.ao_sizes[kr][a2];
    }
    po += nbo;
    L.korep_off[k] = ro; L.korep_size[k] = off; ro += off;
    off = 0;
    for (int b = 0; b < nbx; ++b) {
      int a1 = d->auxrep_irreps[2 * (pp + b)], a2 = d->auxrep_irreps[2 * (pp + b) + 1];
      if (a1 < 0 || a1 >= (int)L.aux_sizes[k].size() || a2 < 0 || a2 >= (int)L.aux_sizes[kr].size()) {
        err = fmt("auxrep block index out of range at k", k, b); return false;
      }
      L.auxrep_irreps[k].push_back({a1, a2}); L.auxrep_offsets[k].push_back(off);
      off += (int64_t)L.aux_sizes[k][a1] * L.aux_sizes[kr][a2];
    }
    pp += nbx;
    L.kauxrep_off[k] = rx; L.kauxrep_size[k] = off; rx += off;
  }
  L.flat_orep = ro; L.flat_auxrep = rx;
  // every k that is not a star representative must have rotation blocks
  for (int k = 0; k < nk; ++k) {
    int ck = L.conj_index[k];
    if (L.index[ck] != ck) {
      if (L.orep_irreps[ck].empty() || L.auxrep_irreps[ck].empty()) {
        err = fmt("k needs rotation blocks but has none", ck); return false;
      }
    } else if (L.irre_list[L.irre_index[ck]] != ck) {
      err = fmt("star representative not at its iBZ slot", ck); return false;
    }
  }

  // V pairs (symmetry_utils.cpp:141-158)
  const size_t np = (size_t)nk * nk;
  L.kpair_off.assign(np, -1); L.kpair_size.assign(np, 0);
  L.V_irreps.resize(np); L.V_offsets.resize(np); L.V_qsize.resize(np);
  int64_t vo = 0, pv = 0;
  for (size_t p = 0; p < np; ++p) {
    int nb = d->kpair_nblocks[p];
    if (nb < 0) { err = fmt("negative kpair_nblocks", (long)p); return false; }
    if (nb == 0) continue;
    int k1 = (int)(p / nk), k2 = (int)(p % nk);
    int64_t off = 0;
    for (int b = 0; b < nb; ++b) {
      int iq = d->V_irreps[3 * (pv + b)], i1 = d->V_irreps[3 * (pv + b) + 1], i2 = d->V_irreps[3 * (pv + b) + 2];
      if (i1 < 0 || i1 >= (int)L.ao_sizes[k1].size() || i2 < 0 || i2 >= (int)L.ao_sizes[k2].size() || iq < 0) {
        err = fmt("V block index out of range at pair", (long)p, b); return false;
      }
      L.V_irreps[p].push_back({iq, i1, i2}); L.V_offsets[p].push_back(off);
      // q size is checked when the pair is bound to a momentum transfer
      off += (int64_t)L.ao_sizes[k1][i1] * L.ao_sizes[k2][i2];  // times q below
    }
    pv += nb;
    L.kpair_off[p] = 0;  // filled after q sizes are known
    L.kpair_size[p] = off;
  }
  // A V block's aux size is that of the pair's momentum transfer.  The P0
  // sweep uses the pair (k, k+q) with aux blocks of q (cugw_qkpt_block.cu:108)
  // and the Sigma sweep (k, k-q) with aux blocks of q (cugw_qkpt_block.cu:170);
  // q(k1,k2) is found from the kpq table.
  std::vector<int> pair_q(np, -1);
  for (int k1 = 0; k1 < nk; ++k1)
    for (int q = 0; q < nk; ++q) {
      int k2 = L.kpq[(size_t)k1 * nk + q];
      size_t p = (size_t)k1 * nk + k2;
      if (pair_q[p] < 0) pair_q[p] = q;
    }
  for (size_t p = 0; p < np; ++p) {
    if (L.V_irreps[p].empty()) continue;
    int q = pair_q[p];
    if (q < 0) { err = fmt("stored pair has no momentum transfer", (long)p); return false; }
    int k1 = (int)(p / nk), k2 = (int)(p % nk);
    int64_t off = 0;
    for (size_t b = 0; b < L.V_irreps[p].size(); ++b) {
      auto &t = L.V_irreps[p][b];
      if (t[0] >= (int)L.aux_sizes[q].size()) { err = fmt("V aux block out of range at pair", (long)p, (long)b); return false; }
      // the Sigma sweep reads the same pair with aux blocks of k1 - k2; they must agree
      int qm = -1;
      for (int qq = 0; qq < nk; ++qq) if (L.kmq[(size_t)k1 * nk + qq] == k2) { qm = qq; break; }
      if (qm >= 0 && L.aux_sizes[qm] != L.aux_sizes[q]) {
        err = fmt("aux blocks of q and -q differ for pair", (long)p); return false;
      }
      L.V_offsets[p][b] = off;
      L.V_qsize[p].push_back(L.aux_sizes[q][t[0]]);
      off += (int64_t)L.aux_sizes[q][t[0]] * L.ao_sizes[k1][t[1]] * L.ao_sizes[k2][t[2]];
    }
    L.kpair_off[p] = vo; L.kpair_size[p] = off; vo += off;
  }
  L.flat_V = vo;
  return true;
}

void flop_model(const Layout &L, double &p0, double &sig, int64_t &nb_p0, int64_t &nb_sig) {
  const int nk = L.nk;
  p0 = 0; sig = 0; nb_p0 = 0; nb_sig = 0;
  for (int iq = 0; iq < L.ink; ++iq) {
    int q = L.irre_list[iq];
    for (int k = 0; k < nk; ++k) {
      int k2 = L.kpq[(size_t)k * nk + q];
      size_t p = (size_t)k * nk + k2;
      for (auto &t : L.V_irreps[p]) {
        double qs = L.aux_sizes[q][t[0]], a = L.ao_sizes[k][t[1]], b = L.ao_sizes[k2][t[2]];
        p0 += L.ns * (L.nts / 2) * 8.0 * (b * qs * a * a + qs * a * b * b + qs * qs * a * b);
        nb_p0++;
      }
    }
  }
  for (int ik = 0; ik < L.ink; ++ik) {
    int k = L.irre_list[ik];
    for (int q = 0; q < nk; ++q) {
      int k2 = L.kmq[(size_t)k * nk + q];
      size_t p = (size_t)k * nk + k2;
      for (auto &t : L.V_irreps[p]) {
        double qs = L.aux_sizes[q][t[0]], a = L.ao_sizes[k][t[1]], b = L.ao_sizes[k2][t[2]];
        sig += L.ns * L.nts * 8.0 * (qs * a * b * b + a * b * qs * qs + a * a * b * qs);
        nb_sig++;
      }
    }
  }
}

namespace {
// odd row stride (in complex elements) keeps the 16 rows an MFMA operand
// fetch touches in distinct LDS banks
int odd_ld(int n) { return n % 2 == 0 ? n + 1 : n; }
// split n rows into the fewest chunks of at most cap rows, balanced
int balanced(int n, int cap) {
  int nch = (n + cap - 1) / cap;
  return (n + nch - 1) / nch;
}
}  // namespace

int64_t pad_to(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Chunking of one Sigma block contraction: qsub aux rows Q by nsub orbitals n
// of k - q per LDS chunk.  Per time of the slab a chunk holds Z
// ([band][qsub*nsub]), E ([band][qsub*nsub]) and conj(P) rows ([qsub][q]).
// Among the shapes that fit, pick the one with the fewest executed MFMA
// FLOPs (16x16x4 tiles, padding included) plus a fixed cost per chunk for
// its barriers and staging.
bool sig_chunks(int q, int k1, int k2, int band, int budget, int TB, double chunk_cost, int max_rows,
                std::map<std::array<int, 4>, std::array<int, 2>> &memo, int &qsub, int &nsub) {
  const std::array<int, 4> key{q, k1, k2, band};
  auto f = memo.find(key);
  if (f != memo.end()) {
    qsub = f->second[0];
    nsub = f->second[1];
    return true;
  }
  // chunk_cost: the default 2.5e5 is ~2000 cycles of a CU's FP64 MFMA rate
  double best = -1;
  for (int qs = 1; qs <= std::min(q, max_rows); ++qs) {
    const int rem = budget - 2 * TB * band;
    if (rem <= 0) break;
    // per orbital of the chunk: Z and E columns of every slab row, and a G column of every orbital m
    const int ns_max = std::min(k2, rem / (2 * TB * band * qs + TB * k2));
    if (ns_max < 1) continue;
    const int qb = balanced(q, qs);
    for (int ns = ns_max; ns >= 1 && ns >= ns_max / 2; --ns) {
      const int nb = balanced(k2, ns);
      double cost = 0;
      for (int Q0 = 0; Q0 < q; Q0 += qb) {
        const int nr = std::min(qb, q - Q0);
        cost += chunk_cost;
        for (int n0 = 0; n0 < k2; n0 += nb) {
          const int nn = std::min(nb, k2 - n0);
          cost += 8.0 * pad_to(TB * nr, 16) * pad_to((int64_t)band * nn, 16) * pad_to(q, 4);     // Z
          cost += 8.0 * pad_to((int64_t)nr * band, 16) * pad_to(TB * nn, 16) * pad_to(k2, 4);  // E
          cost += TB * 8.0 * pad_to(band, 16) * pad_to(band, 16) * pad_to(nr * nn, 4);          // final
          cost += chunk_cost;
        }
      }
      if (best < 0 || cost < best) {
        best = cost;
        qsub = qb;
        nsub = nb;
      }
    }
  }
  if (best < 0) return false;
  memo[key] = {qsub, nsub};
  return true;
}

// (k1, k2) of the V blocks that feed the P0 slabs of aux block iq of q.
static std::vector<std::array<int, 2>> p0_dims(const Layout &L, int q, int iq) {
  std::vector<std::array<int, 2>> d;
  const int nk = L.nk;
  for (int k = 0; k < nk; ++k) {
    const int k2 = L.kpq[(size_t)k * nk + q];
    const size_t p = (size_t)k * nk + k2;
    for (auto &t : L.V_irreps[p])
      if (t[0] == iq) d.push_back({L.ao_sizes[k][t[1]], L.ao_sizes[k2][t[2]]});
  }
  return d;
}

// Rows per P0 slab of an aux block of qs rows.  A chunk of nn orbitals holds E
// and C ([times][band][k1 * nn] each), so narrower bands leave room for wider
// chunks: the E product (columns: times x nn) fills its 16-wide MFMA tiles
// while the final product pads the band to 16 rows.  Among the bands up to
// band_cap, pick the one with the fewest executed MFMA FLOPs (16x16x4 tiles,
// padding included) plus a fixed cost per chunk, summed over the bands of
// the block.
static int p0_band(int qs, int band_cap, int budget, int eb, int TB, double chunk_cost,
                   const std::vector<std::array<int, 2>> &dims) {
  const bool stack = p0_stacked(eb, qs);  // the times share one pass
  const int tb = stack ? TB : 1, passes = stack ? 1 : TB;
  int best_band = balanced(qs, band_cap);
  double best = -1;
  for (int cap = band_cap; cap >= 1; --cap) {
    const int band = balanced(qs, cap);
    const int nbands = (qs + band - 1) / band;
    double cost = 0;
    bool ok = true;
    for (auto &d : dims) {
      const int k1 = d[0], k2 = d[1];
      const int nn_cap = std::min(k2, (budget / (2 * TB * band) - 1) / k1);
      if (nn_cap < 1) { ok = false; break; }
      const int nn = balanced(k2, nn_cap), chunks = (k2 + nn - 1) / nn;
      const double e = 8.0 * pad_to((int64_t)band * k1, 16) * pad_to(tb * nn, 16) * pad_to(k2, 4);
      const double c = tb * 8.0 * pad_to(k1, 16) * pad_to((int64_t)band * nn, 16) * pad_to(k1, 4);
      const double f = tb * 8.0 * 16 * pad_to(qs, 16) * pad_to((int64_t)k1 * nn, 4);
      cost += (double)passes * chunks * (e + c + f + chunk_cost);
    }
    if (!ok) continue;
    cost *= nbands;
    if (best < 0 || cost < best * 0.999) {
      best = cost;
      best_band = band;
    }
  }
  return best_band;
}

// Contiguous block ranges [b, e) of about equal cost, `parts` of them, each
// of at least kSplitMinBlocks blocks (parts <= (end - begin) / kSplitMinBlocks).
static std::vector<int> split_ranges(int begin, int end, int parts, const std::vector<double> &bcost) {
  double tot = 0;
  for (int b = begin; b < end; ++b) tot += bcost[b - begin];
  std::vector<int> cut{begin};
  double acc = 0;
  int b = begin;
  for (int p = 0; p + 1 < parts; ++p) {
    const int left = parts - 1 - p, start = b;
    const double goal = tot * (p + 1) / parts;
    while (b < end - left * kSplitMinBlocks && (b - start < kSplitMinBlocks || acc < goal)) acc += bcost[b++ - begin];
    cut.push_back(b);
  }
  cut.push_back(end);
  return cut;
}

// Launch group of an iBZ point: every point is one group with a resident V
// (one launch covers the sweep); with a streamed V the points of one batch
// (first: [ngroups + 1] first point of each batch).
static int launch_group(const std::vector<int> &first, int ir) {
  if (first.size() < 2) return 0;
  return (int)(std::upper_bound(first.begin(), first.end(), ir) - first.begin()) - 1;
}

// Split the Sigma items over their q where one workgroup per (item, slab)
// would leave the GPU short of work: a system with few iBZ points, times or
// blocks (a sample, a streamed batch of one point), or a few items much
// costlier than the rest (uneven blocks).  Items are costed per slab (8 M N K
// of the Z, E and final products of their blocks) and grouped by launch (one
// slab depth of one streamed batch, or of the whole sweep); an item above
// its group's total / kSplitTarget is cut into that many parts of
// contiguous blocks with balanced cost (at least kSplitMinBlocks blocks
// each).  The parts of an item write partial sums that the reduction kernel
// adds in part order.  The batch prefixes (items of each iBZ point) are
// rebuilt afterwards.
static void split_sigma_items(WorkLists &W, int nts, int ns) {
  const int ngroups = std::max<int>(1, (int)W.sig_group_first.size() - 1);
  W.sig_splits.clear();
  W.sig_part_elems = 0;
  for (int d = 0; d < kNumSigDepths; ++d) {
    const int TB = kSigDepths[d], per = (nts + TB - 1) / TB * ns;
    auto bcost = [&](const SigItem &it, const SigBlk &B) {
      return 8.0 * TB * B.q * B.k2 * ((double)B.q * it.Jb + (double)it.Ib * B.k2 + (double)it.Ib * it.Jb);
    };
    std::vector<double> cost, total(ngroups, 0.0);
    for (const SigItem &it : W.sig_items[d]) {
      double c = 0;
      for (int b = it.blk_begin; b < it.blk_end; ++b) c += bcost(it, W.sig_blks[b]);
      cost.push_back(c);
      total[launch_group(W.sig_group_first, it.ikr)] += c * per;
    }
    std::vector<SigItem> out;
    out.reserve(W.sig_items[d].size());
    for (size_t i = 0; i < W.sig_items[d].size(); ++i) {
      const SigItem it = W.sig_items[d][i];
      const double cap = total[launch_group(W.sig_group_first, it.ikr)] / kSplitTarget;
      const int nb = it.blk_end - it.blk_begin;
      int parts = cap > 0 ? (int)std::ceil(cost[i] / cap) : 1;
      parts = std::max(1, std::min(parts, nb / kSplitMinBlocks));
      if (parts == 1) {
        out.push_back(it);
        continue;
      }
      SigSplit sp;
      sp.part_out = W.sig_part_elems;
      sp.part_stride = (int64_t)per * TB * it.Ib * it.Jb;
      sp.out_base = it.out_base;
      sp.out_tstride = it.out_tstride;
      sp.out_sstride = it.out_sstride;
      sp.nparts = parts;
      sp.TB = TB;
      sp.per = per;
      sp.ns = ns;
      sp.k1 = it.k1;
      sp.J0 = it.J0;
      sp.Ib = it.Ib;
      sp.Jb = it.Jb;
      sp.ikr = it.ikr;
      sp.pad = 0;
      W.sig_splits.push_back(sp);
      W.sig_part_elems += sp.part_stride * parts;
      std::vector<double> bc;
      for (int b = it.blk_begin; b < it.blk_end; ++b) bc.push_back(bcost(it, W.sig_blks[b]));
      const std::vector<int> cut = split_ranges(it.blk_begin, it.blk_end, parts, bc);
      for (int p = 0; p < parts; ++p) {
        SigItem q = it;
        q.blk_begin = cut[p];
        q.blk_end = cut[p + 1];
        q.part_out = sp.part_out + (int64_t)p * sp.part_stride;
        out.push_back(q);
      }
    }
    W.sig_items[d].swap(out);
    // items of iBZ point ikr: [prefix[ikr], prefix[ikr + 1])
    auto &pre = W.sig_batch_items[d];
    const int ink = (int)pre.size();
    pre.assign(ink + 1, 0);
    for (const SigItem &it : W.sig_items[d]) ++pre[it.ikr + 1];
    for (int k = 0; k < ink; ++k) pre[k + 1] += pre[k];
  }
}

// The split of the P0 items over their k, on the rule of split_sigma_items
// (launch groups: the streamed batches of iBZ q, or the whole sweep): an
// item's cost is the work of its blocks for one slab (E, C and the final
// product of TB times of nband rows).
static void split_p0_items(WorkLists &W, int nts, int TB) {
  const int np = (nts / 2 + TB - 1) / TB;
  const int ngroups = std::max<int>(1, (int)W.p0_group_first.size() - 1);
  auto bc = [&](const P0Item &it, const P0Blk &B) {
    return 8.0 * TB * it.nband * ((double)B.k1 * B.k2 * B.k2 + (double)B.k1 * B.k1 * B.k2 + (double)it.q * B.k1 * B.k2);
  };
  std::vector<double> cost, total(ngroups, 0.0);
  for (const P0Item &it : W.p0_items) {
    double c = 0;
    for (int b = it.blk_begin; b < it.blk_end; ++b) c += bc(it, W.p0_blks[b]);
    cost.push_back(c);
    total[launch_group(W.p0_group_first, it.iqr)] += c * np;
  }
  W.p0_splits.clear();
  W.p0_part_elems = 0;
  std::vector<P0Item> out;
  out.reserve(W.p0_items.size());
  for (size_t i = 0; i < W.p0_items.size(); ++i) {
    const P0Item it = W.p0_items[i];
    const double cap = total[launch_group(W.p0_group_first, it.iqr)] / kSplitTarget;
    const int nb = it.blk_end - it.blk_begin;
    int parts = cap > 0 ? (int)std::ceil(cost[i] / cap) : 1;
    parts = std::max(1, std::min(parts, nb / kSplitMinBlocks));
    if (parts == 1) {
      out.push_back(it);
      continue;
    }
    P0Split sp;
    sp.part_out = W.p0_part_elems;
    sp.part_stride = (int64_t)np * TB * it.nband * it.q;
    sp.out_base = it.out_base;
    sp.out_tstride = it.out_tstride;
    sp.nparts = parts;
    sp.TB = TB;
    sp.np = np;
    sp.nband = it.nband;
    sp.q = it.q;
    sp.iqr = it.iqr;
    sp.pad = 0;
    W.p0_splits.push_back(sp);
    W.p0_part_elems += sp.part_stride * parts;
    std::vector<double> bcost;
    for (int b = it.blk_begin; b < it.blk_end; ++b) bcost.push_back(bc(it, W.p0_blks[b]));
    const std::vector<int> cut = split_ranges(it.blk_begin, it.blk_end, parts, bcost);
    for (int p = 0; p < parts; ++p) {
      P0Item q = it;
      q.blk_begin = cut[p];
      q.blk_end = cut[p + 1];
      q.part_out = sp.part_out + (int64_t)p * sp.part_stride;
      out.push_back(q);
    }
  }
  W.p0_items.swap(out);
  auto &pre = W.p0_batch_items;
  const int ink = (int)pre.size();
  pre.assign(ink + 1, 0);
  for (const P0Item &it : W.p0_items) ++pre[it.iqr + 1];
  for (int k = 0; k < ink; ++k) pre[k + 1] += pre[k];
}

bool valid_plan(const Plan &p) {
  return p.sig_chunk_cost >= 0 && p.sig_chunk_cost < 1e12 && p.sig_rows >= 1 && p.sig_rows <= kSigRows &&
         (p.sig_max_times == 1 || p.sig_max_times == 2 || p.sig_max_times == 4) && p.p0_chunk_cost >= 0 &&
         p.p0_chunk_cost < 1e12 && p.p0_band_cap >= 1 && p.p0_band_cap <= kP0Band;
}

bool build_worklists(const Layout &L, int eb, WorkLists &W, std::string &err, bool z8, bool stream,
                     const Plan &plan) {
  W = WorkLists();
  if (!valid_plan(plan)) {
    err = "plan parameters out of range";
    return false;
  }
  W.stream = stream;
  // streamed V: the batch buffer offset of every pair of a batch (pairs of iBZ
  // q: (k, k + q) for P0; of iBZ k: (k, k - q) for Sigma), in k / q order
  std::vector<int64_t> batch_off(stream ? (size_t)L.nk * L.nk : 0, -1);
  int64_t batch_fill = 0;  // complex elements of the open batch
  auto pair_of = [&](int kr, int x, bool p0) {
    return p0 ? (size_t)x * L.nk + L.kpq[(size_t)x * L.nk + kr] : (size_t)kr * L.nk + L.kmq[(size_t)kr * L.nk + x];
  };
  // consecutive iBZ points per batch while their V stays within
  // kStreamBatchBytes (SYMGW_STREAM_BATCH_BYTES overrides it, e.g. 1 for one
  // point per batch)
  int64_t batch_cap = kStreamBatchBytes;
  if (const char *ev = std::getenv("SYMGW_STREAM_BATCH_BYTES")) batch_cap = std::atoll(ev);
  auto groups = [&](bool p0, std::vector<int> &first) {
    first.assign(1, 0);
    int64_t acc = 0;
    for (int ir = 0; ir < L.ink; ++ir) {
      int64_t sz = 0;
      for (int x = 0; x < L.nk; ++x) sz += L.kpair_size[pair_of(L.irre_list[ir], x, p0)];
      if (ir > first.back() && (acc + sz) * eb > batch_cap) {
        first.push_back(ir);
        acc = 0;
      }
      acc += sz;
    }
    first.push_back(L.ink);
  };
  if (stream) {
    groups(true, W.p0_group_first);
    groups(false, W.sig_group_first);
  }
  auto open_batch = [&](std::vector<std::pair<int, int64_t>> &pairs, std::vector<int> &begin) {
    begin.push_back((int)pairs.size());
    batch_fill = 0;
  };
  auto add_point = [&](std::vector<std::pair<int, int64_t>> &pairs, int kr, bool p0) {
    for (int x = 0; x < L.nk; ++x) {
      const size_t p = pair_of(kr, x, p0);
      if (L.V_irreps[p].empty() || batch_off[p] >= 0) continue;  // a pair may recur (k + q == k' + q)
      batch_off[p] = batch_fill;
      pairs.emplace_back((int)p, batch_fill);
      batch_fill += L.kpair_size[p];
    }
    W.batch_V_max = std::max(W.batch_V_max, batch_fill);
  };
  auto close_batch = [&](const std::vector<std::pair<int, int64_t>> &pairs, int from) {
    for (size_t i = from; i < pairs.size(); ++i) batch_off[pairs[i].first] = -1;
  };
  // is iBZ index ir the first / last point of its group
  auto group_starts = [&](const std::vector<int> &first, int ir) {
    return std::binary_search(first.begin(), first.end() - 1, ir);
  };
  auto group_ends = [&](const std::vector<int> &first, int ir) {
    return std::binary_search(first.begin() + 1, first.end(), ir + 1);
  };
  auto v_offset = [&](size_t p, size_t b) {
    return stream ? batch_off[p] + L.V_offsets[p][b] : L.kpair_off[p] + L.V_offsets[p][b];
  };
  const int nk = L.nk, nts = L.nts, ns = L.ns;
  const int budget = (contract_lds(eb) - kDealBytes) / eb;  // complex elements of LDS per workgroup for data
  flop_model(L, W.flop_p0, W.flop_sig, W.n_p0_contractions, W.n_sig_contractions);

  // ---- P0: items per (q_ir, iq, band of <= 16 rows); blocks in (k, V-block) order ----
  size_t p0_lds = 0;
  // kernel mode first: it fixes the times per slab, which the chunking needs
  int64_t n_stacked = 0, n_unstacked = 0;
  for (int iqr = 0; iqr < L.ink; ++iqr)
    for (int qs : L.aux_sizes[L.irre_list[iqr]]) (p0_stacked(eb, qs) ? n_stacked : n_unstacked) += 1;
  W.p0_mode = n_unstacked == 0 ? 1 : n_stacked == 0 ? 0 : 2;
  const int p0_tb = p0_times(eb, W.p0_mode);
  int p0_pair_from = 0;
  for (int iqr = 0; iqr < L.ink; ++iqr) {
    int q = L.irre_list[iqr];
    W.p0_batch_items.push_back((int)W.p0_items.size());
    if (stream) {
      if (group_starts(W.p0_group_first, iqr)) {
        p0_pair_from = (int)W.p0_batch_pairs.size();
        open_batch(W.p0_batch_pairs, W.p0_batch_pair_begin);
      }
      add_point(W.p0_batch_pairs, q, true);
    }
    for (size_t iq = 0; iq < L.aux_sizes[q].size(); ++iq) {
      int qs = L.aux_sizes[q][iq];
      if (qs > 256) { err = fmt("aux block larger than 256", qs); return false; }
      if ((qs + 15) / 16 > contract_waves(eb) * contract_tiles(eb)) {
        err = fmt("P0 slab has more output tiles than the workgroup holds", qs);
        return false;
      }
      // rows per slab: up to kP0Band, fewer where the widest ao block k1 of
      // the slab's V blocks would not leave LDS for one orbital per chunk
      // (E and C rows of k1 + 1 elements for every time and row)
      int maxk1 = 1;
      for (int k = 0; k < nk; ++k) {
        const size_t p = (size_t)k * nk + L.kpq[(size_t)k * nk + q];
        for (auto &t : L.V_irreps[p])
          if (t[0] == (int)iq) maxk1 = std::max(maxk1, L.ao_sizes[k][t[1]]);
      }
      const int band_cap = std::min(plan.p0_band_cap, budget / (2 * p0_tb * (maxk1 + 1)));
      if (band_cap < 1) { err = fmt("P0 ao block too large for the LDS plan (k1)", maxk1); return false; }
      const int band = p0_band(qs, band_cap, budget, eb, p0_tb, plan.p0_chunk_cost, p0_dims(L, q, (int)iq));
      int b0 = (int)W.p0_blks.size();
      for (int k = 0; k < nk; ++k) {
        int k2 = L.kpq[(size_t)k * nk + q];
        size_t p = (size_t)k * nk + k2;
        for (size_t b = 0; b < L.V_irreps[p].size(); ++b) {
          auto &t = L.V_irreps[p][b];
          if (t[0] != (int)iq) continue;
          P0Blk B;
          B.v_off = v_offset(p, b);
          B.k1 = L.ao_sizes[k][t[1]];
          B.k2 = L.ao_sizes[k2][t[2]];
          B.g1_base = L.kao_off[k] * ns * nts + L.ao_offsets[k][t[1]];
          B.g2_base = L.kao_off[k2] * ns * nts + L.ao_offsets[k2][t[2]];
          B.g1_tstride = (int32_t)(ns * L.kao_size[k]); B.g1_sstride = (int32_t)L.kao_size[k];
          B.g2_tstride = (int32_t)(ns * L.kao_size[k2]); B.g2_sstride = (int32_t)L.kao_size[k2];
          // E and C of one chunk ([times][band][k1*nsub] each, +1 element of
          // row padding) share the LDS budget
          const int cap = std::min(B.k2, (budget / (2 * p0_tb * band) - 1) / B.k1);
          if (cap < 1) { err = fmt("P0 block too large for the LDS plan (k1, k2)", B.k1, B.k2); return false; }
          B.nsub = balanced(B.k2, cap);
          B.lde = odd_ld(B.k1 * B.nsub);
          p0_lds = std::max(p0_lds, (size_t)(2 * p0_tb * band * B.lde) * eb);
          W.p0_blks.push_back(B);
        }
      }
      int b1 = (int)W.p0_blks.size();
      for (int Q0 = 0; Q0 < qs; Q0 += band) {
        P0Item it;
        it.blk_begin = b0; it.blk_end = b1;
        it.q = qs; it.Q0 = Q0; it.nband = std::min(band, qs - Q0); it.iqr = iqr;
        it.out_base = (int64_t)nts * L.qaux_off_irre[iqr] + L.aux_offsets[q][iq] + (int64_t)Q0 * qs;
        it.out_tstride = (int32_t)L.qaux_size_irre[iqr];
        it.pad2 = 0;
        it.part_out = -1;
        W.p0_items.push_back(it);
      }
    }
    if (stream && group_ends(W.p0_group_first, iqr)) close_batch(W.p0_batch_pairs, p0_pair_from);
  }
  split_p0_items(W, nts, p0_tb);
  if (stream) W.p0_batch_pair_begin.push_back((int)W.p0_batch_pairs.size());
  // the split-K epilogue passes partial tiles through LDS: one slot per wave of
  // contract_tiles tiles x 64 lanes x 8 reals
  const size_t red_bytes = (size_t)contract_waves(eb) * contract_tiles(eb) * 64 * 8 * (eb / 2);
  // P0 keeps one accumulator per stacked time (at least contract_tiles)
  const size_t red_p0 = (size_t)contract_waves(eb) * std::max(contract_tiles(eb), p0_tb) * 64 * 8 * (eb / 2);
  W.p0_lds = std::max(p0_lds, red_p0) + kDealBytes;

  // ---- Sigma: items per (k_ir, ik1, row band, column band); blocks in (q, V-block) order ----
  size_t sig_lds[kNumSigDepths] = {};
  std::map<std::array<int, 4>, std::array<int, 2>> sig_plans;
  // mixed mode: one INT8 split copy per V block and per P block the Sigma
  // sweep reads, keyed by their FP64 offsets
  std::map<int64_t, size_t> v8_of, p8_of;
  // aux blocks above this size take the INT8 Z product (SYMGW_Z8_MIN_Q
  // overrides the default, e.g. 0 to put every block on it)
  int z8_min_q = kZ8MinQ;
  if (const char *ev = std::getenv("SYMGW_Z8_MIN_Q")) z8_min_q = std::atoi(ev);
  // V split bytes / exponents of the current batch (streamed V) or of the whole V
  int64_t v8_cur = 0, ve_cur = 0;
  int sig_pair_from = 0;
  for (int ikr = 0; ikr < L.ink; ++ikr) {
    int k = L.irre_list[ikr];
    for (int d = 0; d < kNumSigDepths; ++d) W.sig_batch_items[d].push_back((int)W.sig_items[d].size());
    if (stream) {
      if (group_starts(W.sig_group_first, ikr)) {
        sig_pair_from = (int)W.sig_batch_pairs.size();
        open_batch(W.sig_batch_pairs, W.sig_batch_pair_begin);
        if (z8) {
          // batch-relative V offsets repeat from batch to batch: a fresh split buffer
          v8_of.clear();
          v8_cur = ve_cur = 0;
          W.v8_batch_begin.push_back((int)W.v8_split.size());
        }
      }
      add_point(W.sig_batch_pairs, k, false);
    }
    for (size_t i1 = 0; i1 < L.ao_sizes[k].size(); ++i1) {
      int k1 = L.ao_sizes[k][i1];
      SigShape sh = sig_shape(eb, k1);
      sh.times = std::min(sh.times, plan.sig_max_times);
      const int TB = sh.times, band = sh.band;
      int b0 = (int)W.sig_blks.size();
      for (int q = 0; q < nk; ++q) {
        int k2i = L.kmq[(size_t)k * nk + q];
        size_t p = (size_t)k * nk + k2i;
        for (size_t b = 0; b < L.V_irreps[p].size(); ++b) {
          auto &t = L.V_irreps[p][b];
          if (t[1] != (int)i1) continue;
          SigBlk B;
          B.v_off = v_offset(p, b);
          B.q = L.aux_sizes[q][t[0]];
          B.k2 = L.ao_sizes[k2i][t[2]];
          B.g2_base = L.kao_off[k2i] * ns * nts + L.ao_offsets[k2i][t[2]];
          B.g2_tstride = (int32_t)(ns * L.kao_size[k2i]); B.g2_sstride = (int32_t)L.kao_size[k2i];
          B.p_base = L.qaux_off[q] * nts + L.aux_offsets[q][t[0]];
          B.p_tstride = (int32_t)L.qaux_size[q];
          if (!sig_chunks(B.q, k1, B.k2, band, budget, TB, plan.sig_chunk_cost, plan.sig_rows, sig_plans, B.qsub,
                          B.nsub)) {
            err = fmt("Sigma block too large for the LDS plan (q, k2)", B.q, B.k2);
            return false;
          }
          B.ldz = odd_ld(B.qsub * B.nsub);
          B.lde = odd_ld(B.qsub * B.nsub);
          B.kp = 0;
          B.v8_off = B.ve_off = B.p8_off = B.pe_off = 0;
          if (z8 && B.q > z8_min_q) {
            B.kp = (B.q + 15) / 16 * 16;
            const int ncol = k1 * B.k2;
            auto iv = v8_of.find(B.v_off);
            if (iv == v8_of.end()) {
              // B operand of Z: columns (a, b) of V, K = P (stride k1 k2)
              Split8Blk S8{B.v_off, 0, v8_cur, ve_cur, 1, ncol, ncol, B.q, B.kp, 1, 1, 0};
              iv = v8_of.emplace(B.v_off, W.v8_split.size()).first;
              W.v8_split.push_back(S8);
              v8_cur += (int64_t)kZ8Comps * kZ8Slices * ncol * B.kp;
              ve_cur += (int64_t)kZ8Comps * ncol;
              W.v8_bytes = std::max(W.v8_bytes, v8_cur);
              W.ve_count = std::max(W.ve_count, ve_cur);
            }
            B.v8_off = W.v8_split[iv->second].dst_off;
            B.ve_off = W.v8_split[iv->second].e_off;
            auto ip = p8_of.find(B.p_base);
            if (ip == p8_of.end()) {
              // A operand of Z: rows (t, Q) of P, K = P (contiguous)
              Split8Blk S8{B.p_base, B.p_tstride, W.p8_bytes, W.pe_count, B.q, 1, B.q, B.q, B.kp, nts, 0, 0};
              ip = p8_of.emplace(B.p_base, W.p8_split.size()).first;
              W.p8_split.push_back(S8);
              W.p8_bytes += (int64_t)nts * kZ8Comps * kZ8Slices * B.q * B.kp;
              W.pe_count += (int64_t)nts * kZ8Comps * B.q;
            }
            B.p8_off = W.p8_split[ip->second].dst_off;
            B.pe_off = W.p8_split[ip->second].e_off;
          }
          size_t &lds_t = sig_lds[sig_depth_index(TB)];
          lds_t = std::max(lds_t, (size_t)TB * (band * (B.ldz + B.lde) + B.k2 * B.nsub) * eb);
          W.sig_blks.push_back(B);
        }
      }
      int b1 = (int)W.sig_blks.size();
      // blocks in ascending LDS row stride, so that the Z and E buffers never
      // shrink from one block to the next and the kernel's block-boundary
      // wait (the next G chunk must not land on buffers still being read) is
      // never needed; the order is the same for every slab of the item
      std::stable_sort(W.sig_blks.begin() + b0, W.sig_blks.begin() + b1,
                       [](const SigBlk &x, const SigBlk &y) { return x.ldz + x.lde < y.ldz + y.lde; });
      {
        const int nt = (band + 15) / 16;
        if (TB * nt * nt > contract_waves(eb) * contract_tiles(eb)) {
          err = fmt("Sigma slab has more output tiles than the workgroup holds", band);
          return false;
        }
      }
      for (int I0 = 0; I0 < k1; I0 += band)
        for (int J0 = 0; J0 < k1; J0 += band) {
          SigItem it;
          it.blk_begin = b0; it.blk_end = b1;
          it.k1 = k1; it.I0 = I0; it.Ib = std::min(band, k1 - I0);
          it.J0 = J0; it.Jb = std::min(band, k1 - J0); it.ikr = ikr;
          it.part_out = -1;
          it.out_base = L.kao_off_irre[ikr] * ns * nts + L.ao_offsets[k][i1] + (int64_t)I0 * k1;
          it.out_tstride = (int32_t)(ns * L.kao_size_irre[ikr]);
          it.out_sstride = (int32_t)L.kao_size_irre[ikr];
          W.sig_items[sig_depth_index(TB)].push_back(it);
        }
    }
    if (stream && group_ends(W.sig_group_first, ikr)) close_batch(W.sig_batch_pairs, sig_pair_from);
  }
  split_sigma_items(W, nts, ns);
  if (stream) W.sig_batch_pair_begin.push_back((int)W.sig_batch_pairs.size());
  if (stream && z8) W.v8_batch_begin.push_back((int)W.v8_split.size());
  // the INT8-Z Sigma kernel reduces split-K parts over its own wave / tile shape
  const size_t red_sig = z8 ? (size_t)kZ8Waves * kZ8Tiles * 64 * 8 * (eb / 2) : red_bytes;
  for (int d = 0; d < kNumSigDepths; ++d)
    W.sig_lds[d] = W.sig_items[d].empty() ? 0 : std::max(sig_lds[d], red_sig) + kDealBytes;
  for (int d = 0; d < kNumSigDepths; ++d)
    if (W.sig_lds[d] > (size_t)contract_lds(eb)) { err = fmt("Sigma LDS plan exceeds the CU's LDS (bytes)", (int)W.sig_lds[d]); return false; }
  if (W.p0_lds > (size_t)contract_lds(eb)) { err = fmt("P0 LDS plan exceeds the CU's LDS (bytes)", (int)W.p0_lds); return false; }

  // ---- rotations of G (ao) and P (aux) to the full BZ ----
  for (int which = 0; which < 2; ++which) {
    auto &out = which == 0 ? W.g_rot : W.p_rot;
    for (int k = 0; k < nk; ++k) {
      int ck = L.conj_index[k];
      int ick = L.irre_index[ck];
      int kr = L.irre_list[ick];
      const auto &sz = which == 0 ? L.ao_sizes[ck] : L.aux_sizes[ck];
      const auto &so = which == 0 ? L.ao_offsets[ck] : L.aux_offsets[ck];
      const auto &szr = which == 0 ? L.ao_sizes[kr] : L.aux_sizes[kr];
      const auto &sor = which == 0 ? L.ao_offsets[kr] : L.aux_offsets[kr];
      if (L.index[ck] == ck) {
        for (size_t b = 0; b < sz.size(); ++b) {
          RotBlk R{so[b], sor[b], -1, sz[b], szr[b], k, k != ck ? 1 : 0};
          out.push_back(R);
        }
      } else {
        const auto &irr = which == 0 ? L.orep_irreps[ck] : L.auxrep_irreps[ck];
        const auto &off = which == 0 ? L.orep_offsets[ck] : L.auxrep_offsets[ck];
        int64_t base = which == 0 ? L.korep_off[ck] : L.kauxrep_off[ck];
        for (size_t b = 0; b < irr.size(); ++b) {
          int a1 = irr[b][0], a2 = irr[b][1];
          // the rotation kernel keeps a column chunk of at least 16 columns of
          // X O^H (n rows) in its LDS
          if (szr[a2] > kRotMaxBlock) {
            err = fmt(which == 0 ? "rotated ao block larger than 256" : "rotated aux block larger than 256", szr[a2]);
            return false;
          }
          RotBlk R{so[a1], sor[a2], base + off[b], sz[a1], szr[a2], k, k != ck ? 1 : 0};
          out.push_back(R);
        }
      }
    }
  }
  // ---- W and Dyson solve blocks ----
  for (int iqr = 0; iqr < L.ink; ++iqr) {
    int q = L.irre_list[iqr];
    for (size_t b = 0; b < L.aux_sizes[q].size(); ++b) {
      int n = L.aux_sizes[q][b];
      W.w_blocks.push_back(SolveBlk{L.aux_offsets[q][b], W.w_scr_group, n, iqr});
      W.w_scr_group += 2 * (int64_t)n * n;
    }
  }
  for (int ikr = 0; ikr < L.ink; ++ikr) {
    int k = L.irre_list[ikr];
    for (size_t b = 0; b < L.ao_sizes[k].size(); ++b) {
      int n = L.ao_sizes[k][b];
      W.d_blocks.push_back(SolveBlk{L.ao_offsets[k][b], W.d_scr_group, n, ikr});
      W.d_scr_group += (int64_t)n * n;
    }
  }
  return true;
}

}  // namespace symgw
