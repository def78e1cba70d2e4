"""Block-sparse GW systems: the layout tables a symgw context is built from,
plus the host inputs (rotation blocks, F/S, transforms).

Two sources:
  * the reference's LiH 2x2x2 fixture (converted to tests/golden/lih_nk2.npz),
    read as symmetry_utils_t reads symmetry_info.h5
    (SymmetryMBPT/src/common/symmetry_utils.cpp:96-158) and as
    sc_loop_block_t::read_input_data reads input_flat.h5
    (SymmetryMBPT/src/scmbpt/sc_loop_block_t.cpp:64-109);
  * seeded synthetic systems (the benchmark configurations): Gamma-centred
    n^3 mesh, cubic-star irreducible set, Haar-random block rotations, the
    abelian V-block rule ik2 = (ik1 + iq) mod nblocks(k2), Hermitian Fock
    matrices U diag(eps) U^H with eps ~ U[-1, 1] Ha, S = I.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from .transforms import Transforms


def haar_unitary(rng: np.random.Generator, n: int) -> np.ndarray:
    """Haar-random unitary: QR of a complex Ginibre matrix with the phase fix
    of Mezzadri, Notices AMS 54 (2007) 592."""
    z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))[None, :]


@dataclass
class System:
    nk: int
    ink: int
    ns: int
    transforms: Transforms
    ao_sizes: List[List[int]]           # per full-BZ k
    aux_sizes: List[List[int]]          # per full-BZ q
    irre_list: np.ndarray
    index: np.ndarray
    irre_index: np.ndarray
    conj_index: np.ndarray
    kpq: np.ndarray                     # (nk, nk)
    kmq: np.ndarray
    orep_irreps: List[List[Tuple[int, int]]]
    auxrep_irreps: List[List[Tuple[int, int]]]
    V_irreps: Dict[int, List[Tuple[int, int, int]]]  # pair index -> blocks
    kweights: np.ndarray                # (ink,) star sizes, summing to nk
    orep: np.ndarray = None             # flat complex
    auxrep: np.ndarray = None
    F: np.ndarray = None                # (ns, flat_ao) complex, iBZ
    S: np.ndarray = None
    H: np.ndarray = None
    mu: float = 0.0
    name: str = ""
    v_seed: int = 0
    v_scale: float = 1.0
    V: Optional[np.ndarray] = None      # host V (fixture systems)
    _keep: list = field(default_factory=list, repr=False)

    # ---- sizes (own cumulative sums) ----
    @property
    def nts(self):
        return self.transforms.nts

    def kslice(self, k):
        return int(sum(b * b for b in self.ao_sizes[k]))

    def qslice(self, q):
        return int(sum(b * b for b in self.aux_sizes[q]))

    @property
    def flat_ao(self):
        return sum(self.kslice(int(k)) for k in self.irre_list)

    @property
    def flat_aux(self):
        return sum(self.qslice(int(q)) for q in self.irre_list)

    @property
    def flat_ao_full(self):
        return sum(self.kslice(k) for k in range(self.nk))

    def kao_off_irre(self):
        out, o = [], 0
        for k in self.irre_list:
            out.append(o)
            o += self.kslice(int(k))
        return out

    def qaux_off_irre(self):
        out, o = [], 0
        for q in self.irre_list:
            out.append(o)
            o += self.qslice(int(q))
        return out

    def ao_block_offsets(self, k):
        return list(np.concatenate([[0], np.cumsum([b * b for b in self.ao_sizes[k]])[:-1]]).astype(int))

    def aux_block_offsets(self, q):
        return list(np.concatenate([[0], np.cumsum([b * b for b in self.aux_sizes[q]])[:-1]]).astype(int))

    def pair_q(self, k1, k2):
        hit = np.nonzero(self.kpq[k1] == k2)[0]
        return int(hit[0])

    def V_block_sizes(self, p):
        k1, k2 = divmod(p, self.nk)
        q = self.pair_q(k1, k2)
        return [(self.aux_sizes[q][iq], self.ao_sizes[k1][i1], self.ao_sizes[k2][i2])
                for iq, i1, i2 in self.V_irreps.get(p, [])]

    def kpair_offsets(self):
        """pair -> (offset, size) of the flat V slice, in pair order."""
        out, o = {}, 0
        for p in range(self.nk * self.nk):
            if p in self.V_irreps and self.V_irreps[p]:
                sz = sum(a * b * c for a, b, c in self.V_block_sizes(p))
                out[p] = (o, sz)
                o += sz
        return out

    def flat_V(self):
        return sum(sz for _, sz in self.kpair_offsets().values())

    # ---- C descriptor ----
    def desc(self) -> _lib.LayoutDesc:
        nk = self.nk
        arrs = {}

        def a32(x):
            v = np.ascontiguousarray(np.asarray(x, dtype=np.int32).ravel())
            self._keep.append(v)
            return v.ctypes.data_as(C.POINTER(C.c_int32))

        d = _lib.LayoutDesc()
        d.nk, d.ink, d.ns, d.nts = nk, self.ink, self.ns, self.nts
        d.nw, d.nw_b = self.transforms.nw, self.transforms.nw_b
        d.kao_nblocks = a32([len(s) for s in self.ao_sizes])
        d.ao_sizes = a32([b for s in self.ao_sizes for b in s])
        d.qaux_nblocks = a32([len(s) for s in self.aux_sizes])
        d.aux_sizes = a32([b for s in self.aux_sizes for b in s])
        d.irre_list = a32(self.irre_list)
        d.index = a32(self.index)
        d.irre_index = a32(self.irre_index)
        d.conj_index = a32(self.conj_index)
        d.kpq = a32(self.kpq)
        d.kmq = a32(self.kmq)
        d.orep_nblocks = a32([len(x) for x in self.orep_irreps])
        d.orep_irreps = a32([v for x in self.orep_irreps for pr in x for v in pr] or [0])
        d.auxrep_nblocks = a32([len(x) for x in self.auxrep_irreps])
        d.auxrep_irreps = a32([v for x in self.auxrep_irreps for pr in x for v in pr] or [0])
        d.kpair_nblocks = a32([len(self.V_irreps.get(p, [])) for p in range(nk * nk)])
        d.V_irreps = a32([v for p in range(nk * nk) for t in self.V_irreps.get(p, []) for v in t] or [0])
        del arrs
        return d

    # ---- host-side arrays in the reference layouts ----
    def zeros_ts_kij(self):
        return np.zeros((self.nts, self.ns, self.flat_ao), dtype=np.complex128)

    def zeros_q_t_QP(self):
        return np.zeros(self.nts * self.flat_aux, dtype=np.complex128)

    def host_V(self, pairs=None):
        """Flat V over the stored pairs (fixture V, or the host generator twin)."""
        if self.V is not None:
            return self.V
        offs = self.kpair_offsets()
        out = np.zeros(self.flat_V(), dtype=np.complex128)
        sel = [p for p in offs if pairs is None or p in pairs]
        Vs, st = _lib.generate_V_pairs_host(self, self.v_seed, self.v_scale, self.decay_table(),
                                            [divmod(p, self.nk) for p in sel])
        for p, o in zip(sel, st):
            out[offs[p][0]:offs[p][0] + offs[p][1]] = Vs[o:o + offs[p][1]]
        return out

    def host_V_pairs(self, pairs):
        """V of the listed stored pairs only: (compact array, {pair: offset})."""
        sel = sorted(set(int(p) for p in pairs))
        Vs, st = _lib.generate_V_pairs_host(self, self.v_seed, self.v_scale, self.decay_table(),
                                            [divmod(p, self.nk) for p in sel])
        return Vs, {p: int(o) for p, o in zip(sel, st)}

    def decay_table(self):
        nao = max(sum(s) for s in self.ao_sizes)
        return np.exp(-np.arange(nao + 1) / 4.0)


# ---------------------------------------------------------------------------
# k meshes
# ---------------------------------------------------------------------------
def cubic_mesh(n: int):
    """Gamma-centred n^3 mesh: index i*n^2 + j*n + l; kpq/kmq tables."""
    pts = np.array([(i, j, l) for i in range(n) for j in range(n) for l in range(n)], dtype=np.int64)
    nk = len(pts)

    def idx(v):
        v = np.mod(v, n)
        return v[..., 0] * n * n + v[..., 1] * n + v[..., 2]

    kpq = idx(pts[:, None, :] + pts[None, :, :]).astype(np.int32)
    kmq = idx(pts[:, None, :] - pts[None, :, :]).astype(np.int32)
    return pts, kpq.reshape(nk, nk), kmq.reshape(nk, nk)


def cubic_stars(pts: np.ndarray, n: int, ink_target: Optional[int] = None):
    """Stars of the cubic point group O_h on the mesh (key: sorted folded
    coordinates).  When ink_target is smaller than the number of O_h stars the
    surplus stars are merged into earlier ones (a synthetic reduction)."""
    fold = np.minimum(pts, n - pts)
    keys = [tuple(sorted(f, reverse=True)) for f in fold]
    order = {}
    for i, kk in enumerate(keys):
        order.setdefault(kk, []).append(i)
    stars = [v for _, v in sorted(order.items(), key=lambda kv: kv[1][0])]
    if ink_target is not None and ink_target < len(stars):
        extra = stars[ink_target:]
        stars = stars[:ink_target]
        for j, st in enumerate(extra):
            stars[1 + j % (ink_target - 1)].extend(st)
    return stars


def split_sizes(rng, total: int, nblocks: int, lo: int, hi: int) -> List[int]:
    """nblocks sizes log-uniform in [lo, hi], rescaled to sum to total (each
    kept inside [1, max(hi, total)])."""
    for _ in range(1000):
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), nblocks))
        s = np.maximum(1, np.round(x * total / x.sum())).astype(int)
        d = total - s.sum()
        i = 0
        while d != 0:
            j = i % nblocks
            if d > 0:
                s[j] += 1; d -= 1
            elif s[j] > 1:
                s[j] -= 1; d += 1
            i += 1
        if s.min() >= 1:
            return [int(v) for v in s]
    raise RuntimeError("could not split sizes")


def synthetic(name: str, n: int, ink: int, ao_sizes: List[int], aux_sizes: List[int], ns: int,
              transforms: Transforms, seed: int, v_scale: float = 1.0, all_pairs: bool = False) -> System:
    """Seeded synthetic system with the same block structure at every k."""
    rng = np.random.default_rng(seed)
    pts, kpq, kmq = cubic_mesh(n)
    nk = len(pts)
    stars = cubic_stars(pts, n, ink)
    ink = len(stars)
    irre_list = np.array([st[0] for st in stars], dtype=np.int32)
    index = np.zeros(nk, dtype=np.int32)
    irre_index = np.zeros(nk, dtype=np.int32)
    for i, st in enumerate(stars):
        for k in st:
            index[k] = st[0]
            irre_index[k] = i
    conj_index = np.arange(nk, dtype=np.int32)
    kweights = np.array([len(st) for st in stars], dtype=np.float64)
    return _assemble(name, rng, kpq, kmq, irre_list, index, irre_index, conj_index, kweights, ao_sizes, aux_sizes,
                     ns, transforms, seed, v_scale, all_pairs)


def time_reversal_square(ns: int = 1, ao=(3, 2), aux=(4, 3), ni: int = 8, beta: float = 10.0, seed: int = 13,
                         all_pairs: bool = False) -> System:
    """3x3 two-dimensional mesh whose symmetry is the x <-> y mirror plus time
    reversal k -> -k.  iBZ = {(0,0), (1,0), (1,1), (1,2)}; the other points
    cover every way the reference derives a full-BZ slice
    (cugw_routines_block.cu:124-186): rotation of a representative ((0,1),
    (2,1)), complex conjugate of a representative ((2,0), (2,2)), and complex
    conjugate of a rotated point ((0,2) = -(0,1)).  Every k carries the same
    block sizes and each rotation maps block b to block b."""
    from .transforms import chebyshev
    n = 3
    pts = np.array([(i, j) for i in range(n) for j in range(n)], dtype=np.int64)
    nk = len(pts)

    def idx(v):
        v = np.mod(v, n)
        return int(v[0] * n + v[1])

    kpq = np.array([[idx(a + b) for b in pts] for a in pts], dtype=np.int32)
    kmq = np.array([[idx(a - b) for b in pts] for a in pts], dtype=np.int32)
    reps = [idx((0, 0)), idx((1, 0)), idx((1, 1)), idx((1, 2))]
    irre_list = np.array(reps, dtype=np.int32)
    index = np.zeros(nk, dtype=np.int32)
    conj_index = np.arange(nk, dtype=np.int32)
    for k, v in enumerate(pts):
        mir = idx(v[::-1])
        if k in reps:
            index[k] = k
        elif mir in reps:                     # reached by the mirror
            index[k] = mir
        else:                                 # reached by time reversal (after the mirror or not)
            ck = idx(-v)
            conj_index[k] = ck
            index[k] = ck if ck in reps else idx(pts[ck][::-1])
            assert index[k] in reps
    for k in range(nk):                       # the TR partner's own representative
        ck = conj_index[k]
        if ck != k:
            assert conj_index[ck] == ck
    irre_index = np.array([reps.index(int(index[k])) for k in range(nk)], dtype=np.int32)
    kweights = np.array([float(np.sum(irre_index == i)) for i in range(len(reps))])
    rng = np.random.default_rng(seed)
    return _assemble("tr_square", rng, kpq, kmq, irre_list, index, irre_index, conj_index, kweights, list(ao),
                     list(aux), ns, chebyshev(ni, beta), seed, 1.0, all_pairs)


def permuted_blocks(n: int = 3, ns: int = 1, ni: int = 8, beta: float = 10.0, seed: int = 17,
                    max_blocks: int = 4, max_size: int = 9) -> System:
    """Rotations with permuted block maps and k-dependent block structure: on
    the n^3 mesh with its O_h stars, every star has its own irreps (2 to
    max_blocks orbital and auxiliary blocks of sizes 1..max_size), and every
    non-representative point k holds its representative's blocks in its own
    order: block b of k is block sigma_k(b) of the representative, so the
    rotation maps (orep_irreps / auxrep_irreps, cugw_routines_block.cu:134-186)
    pair blocks a1 = b with a2 = sigma_k(b) != b.  The auxiliary order of q and
    -q is the same (P(-q) = P(q)^* reads one layout).  V blocks of pair
    (k1, k2) with transfer q: aux block iq couples ao block iq mod n(k1) of k1
    to (iq + iq mod n(k1)) mod n(k2) of k2."""
    from .transforms import chebyshev
    rng = np.random.default_rng(seed)
    pts, kpq, kmq = cubic_mesh(n)
    nk = len(pts)
    stars = cubic_stars(pts, n)
    ink = len(stars)
    irre_list = np.array([st[0] for st in stars], dtype=np.int32)
    index = np.zeros(nk, dtype=np.int32)
    irre_index = np.zeros(nk, dtype=np.int32)
    for i, st in enumerate(stars):
        for k in st:
            index[k] = st[0]
            irre_index[k] = i

    def sizes():
        nb = int(rng.integers(2, max_blocks + 1))
        return [int(v) for v in rng.integers(1, max_size + 1, nb)]
    rep_ao = {int(r): sizes() for r in irre_list}
    rep_aux = {int(r): sizes() for r in irre_list}
    ao, aux = [None] * nk, [None] * nk
    omap, amap = [[] for _ in range(nk)], [[] for _ in range(nk)]
    for k in range(nk):
        r = int(index[k])
        if k == r:
            ao[k], aux[k] = list(rep_ao[r]), list(rep_aux[r])
            continue
        so = rng.permutation(len(rep_ao[r]))
        ao[k] = [rep_ao[r][int(s)] for s in so]
        omap[k] = [(b, int(so[b])) for b in range(len(so))]
    for q in range(nk):
        r = int(index[q])
        if q == r or aux[q] is not None:
            continue
        mq = int(kmq[0, q])
        sa = rng.permutation(len(rep_aux[r]))
        for x in {q, mq}:
            if x != int(index[x]):
                aux[x] = [rep_aux[r][int(s)] for s in sa]
                amap[x] = [(b, int(sa[b])) for b in range(len(sa))]
    for q in range(nk):                     # q and -q: one auxiliary order
        mq = int(kmq[0, q])
        assert aux[q] == aux[mq], (q, mq)
    orep, auxrep = [], []
    for k in range(nk):
        for a1, _ in omap[k]:
            orep.append(haar_unitary(rng, ao[k][a1]).ravel())
        for a1, _ in amap[k]:
            auxrep.append(haar_unitary(rng, aux[k][a1]).ravel())
    need = set()
    for q in irre_list:
        for k in range(nk):
            need.add(k * nk + int(kpq[k, q]))
    for k in irre_list:
        for q in range(nk):
            need.add(int(k) * nk + int(kmq[k, q]))
    V_irreps = {}
    for p in sorted(need):
        k1, k2 = divmod(p, nk)
        q = int(np.nonzero(kpq[k1] == k2)[0][0])
        n1, n2 = len(ao[k1]), len(ao[k2])
        V_irreps[p] = [(iq, iq % n1, (iq + iq % n1) % n2) for iq in range(len(aux[q]))]
    F, S = [], []
    for s in range(ns):
        fs, ss = [], []
        for r in irre_list:
            for b in ao[int(r)]:
                U = haar_unitary(rng, b)
                fs.append((U @ np.diag(rng.uniform(-1.0, 1.0, b)) @ U.conj().T).ravel())
                ss.append(np.eye(b).ravel())
        F.append(np.concatenate(fs))
        S.append(np.concatenate(ss))
    F = np.array(F)
    S = np.array(S).astype(np.complex128)
    kweights = np.array([len(st) for st in stars], dtype=np.float64)
    return System(nk=nk, ink=ink, ns=ns, transforms=chebyshev(ni, beta), ao_sizes=ao, aux_sizes=aux,
                  irre_list=irre_list, index=index, irre_index=irre_index, conj_index=np.arange(nk, dtype=np.int32),
                  kpq=kpq, kmq=kmq, orep_irreps=omap, auxrep_irreps=amap, V_irreps=V_irreps, kweights=kweights,
                  orep=np.concatenate(orep) if orep else np.zeros(0, np.complex128),
                  auxrep=np.concatenate(auxrep) if auxrep else np.zeros(0, np.complex128),
                  F=F, S=S, H=np.zeros_like(F), name="permuted_blocks", v_seed=seed,
                  v_scale=1.0 / float(np.sqrt(max(sum(a) for a in aux))))


def _assemble(name, rng, kpq, kmq, irre_list, index, irre_index, conj_index, kweights, ao_sizes, aux_sizes, ns,
              transforms, seed, v_scale, all_pairs) -> System:
    """Rotation blocks, V pair list, Fock matrices and the System record of a
    mesh whose points all carry the same block sizes."""
    nk = len(index)
    ink = len(irre_list)
    aos = [list(ao_sizes) for _ in range(nk)]
    auxs = [list(aux_sizes) for _ in range(nk)]
    orep_irreps, auxrep_irreps, orep, auxrep = [], [], [], []
    for k in range(nk):
        if index[k] == k or conj_index[k] != k:   # representative, or derived through its TR partner
            orep_irreps.append([])
            auxrep_irreps.append([])
            continue
        orep_irreps.append([(b, b) for b in range(len(ao_sizes))])
        auxrep_irreps.append([(b, b) for b in range(len(aux_sizes))])
        for b in ao_sizes:
            orep.append(haar_unitary(rng, b).ravel())
        for b in aux_sizes:
            auxrep.append(haar_unitary(rng, b).ravel())
    # V pairs: everything the P0 and Sigma sweeps read
    need = set()
    for q in irre_list:
        for k in range(nk):
            need.add(k * nk + int(kpq[k, q]))
    for k in irre_list:
        for q in range(nk):
            need.add(int(k) * nk + int(kmq[k, q]))
    if all_pairs:
        need = set(range(nk * nk))
    nb_ao, nb_aux = len(ao_sizes), len(aux_sizes)
    V_irreps = {p: [(iq, i1, (i1 + iq) % nb_ao) for iq in range(nb_aux) for i1 in range(nb_ao)] for p in sorted(need)}
    # Fock matrices of the iBZ points, S = I, H = 0
    nblk = len(ao_sizes)
    F = np.zeros((ns, sum(b * b for b in ao_sizes) * ink), dtype=np.complex128)
    S = np.zeros_like(F)
    for s in range(ns):
        o = 0
        for i in range(ink):
            for b in ao_sizes:
                U = haar_unitary(rng, b)
                eps = rng.uniform(-1.0, 1.0, b)
                F[s, o:o + b * b] = (U @ np.diag(eps) @ U.conj().T).ravel()
                S[s, o:o + b * b] = np.eye(b).ravel()
                o += b * b
    sysm = System(nk=nk, ink=ink, ns=ns, transforms=transforms, ao_sizes=aos, aux_sizes=auxs,
                  irre_list=irre_list, index=index, irre_index=irre_index, conj_index=conj_index,
                  kpq=kpq, kmq=kmq, orep_irreps=orep_irreps, auxrep_irreps=auxrep_irreps, V_irreps=V_irreps,
                  kweights=kweights,
                  orep=np.concatenate(orep) if orep else np.zeros(0, np.complex128),
                  auxrep=np.concatenate(auxrep) if auxrep else np.zeros(0, np.complex128),
                  F=F, S=S, H=np.zeros_like(F), name=name, v_seed=seed, v_scale=v_scale)
    del nblk
    return sysm


def lih_fixture(z, transforms: Transforms) -> System:
    """The reference's LiH 2x2x2 test system (nk = ink = 8, ns = 2, nao = 11,
    NQ = 58), no rotations (SymmetryMBPT/test/test_cugw.cpp:19-34)."""
    nk = 8
    kmesh = z["meta__grid__k_mesh_scaled"]
    n = 2
    ijk = np.round(kmesh * n).astype(int) % n
    lin = ijk[:, 0] * n * n + ijk[:, 1] * n + ijk[:, 2]
    # map mesh coordinates back to the fixture's k order
    pos = {int(l): i for i, l in enumerate(lin)}
    kpq = np.zeros((nk, nk), dtype=np.int32)
    kmq = np.zeros((nk, nk), dtype=np.int32)
    for a in range(nk):
        for b in range(nk):
            s = (ijk[a] + ijk[b]) % n
            d = (ijk[a] - ijk[b]) % n
            kpq[a, b] = pos[int(s[0] * n * n + s[1] * n + s[2])]
            kmq[a, b] = pos[int(d[0] * n * n + d[1] * n + d[2])]
    aos = [list(z[f"sym__block__ao_sizes__{k}"].astype(int)) for k in range(nk)]
    auxs = [list(z[f"sym__block__aux_sizes__{k}"].astype(int)) for k in range(nk)]
    V_irreps = {}
    for p in range(nk * nk):
        t = z[f"sym__block__tensor_irreps__{p}"].reshape(-1, 3)
        V_irreps[p] = [tuple(int(v) for v in row) for row in t]
    ar = np.arange(nk, dtype=np.int32)
    F = z["flat__Fock-k"][..., 0] + 1j * z["flat__Fock-k"][..., 1]
    S = z["flat__S-k"][..., 0] + 1j * z["flat__S-k"][..., 1]
    H = z["flat__H-k"][..., 0] + 1j * z["flat__H-k"][..., 1]
    sysm = System(nk=nk, ink=nk, ns=2, transforms=transforms, ao_sizes=aos, aux_sizes=auxs, irre_list=ar.copy(),
                  index=ar.copy(), irre_index=ar.copy(), conj_index=ar.copy(), kpq=kpq, kmq=kmq,
                  orep_irreps=[[] for _ in range(nk)], auxrep_irreps=[[] for _ in range(nk)], V_irreps=V_irreps,
                  kweights=np.ones(nk), orep=np.zeros(0, np.complex128), auxrep=np.zeros(0, np.complex128),
                  F=F, S=S, H=H, name="LiH-nk2")
    # make_hermitian_kao (sc_loop_block_t.cpp:101-103): per block (X + X^H)/2
    for arr in (sysm.F, sysm.S, sysm.H):
        for s in range(2):
            o = 0
            for k in range(nk):
                for b in aos[k]:
                    m = arr[s, o:o + b * b].reshape(b, b)
                    arr[s, o:o + b * b] = (0.5 * (m + m.conj().T)).ravel()
                    o += b * b
    V = z["V_flat"]
    sysm.V = np.ascontiguousarray(V[0::2] + 1j * V[1::2])
    return sysm


# ---------------------------------------------------------------------------
# named configurations
# ---------------------------------------------------------------------------
SEED_BASE = 20241203


def coulomb_scale(aux_sizes) -> float:
    """V scale of the benchmark systems: 1/sqrt(naux), so that the bare
    Coulomb integrals (ij|kl) = sum_Q V^Q_ij conj(V^Q_kl) of unit-variance
    entries are O(1) Ha (unscaled entries make them O(naux) Ha, a self-energy
    of hundreds of Ha and a Dyson step no imaginary-time grid resolves)."""
    return 1.0 / float(np.sqrt(sum(aux_sizes)))


def config1(transforms: Optional[Transforms] = None) -> System:
    """2x2x2 mesh, 4 iBZ points, nao 26 {8,6,6,4,2}, naux 128 {40,32,32,24},
    beta = 50, 60-node Chebyshev grid, ns = 1."""
    from .transforms import chebyshev
    tr = transforms or chebyshev(60, 50.0)
    aux = [40, 32, 32, 24]
    return synthetic("config1", 2, 4, [8, 6, 6, 4, 2], aux, 1, tr, SEED_BASE + 1, v_scale=coulomb_scale(aux),
                     all_pairs=True)


def config2_sizes():
    rng = np.random.default_rng(SEED_BASE + 2)
    ao = split_sizes(rng, 104, 8, 4, 32)
    aux = split_sizes(rng, 400, 8, 16, 128)
    return ao, aux


def config2(transforms: Optional[Transforms] = None, v_scale: Optional[float] = None) -> System:
    """6x6x6 mesh reduced to 16 iBZ points; nao 104 in 8 blocks log-uniform
    [4, 32]; naux 400 in 8 blocks log-uniform [16, 128]; beta = 100; IR grid
    of cutoff Lambda = 1e4 with 104 fermionic (nts = 106) and 103 bosonic
    points, generated by the library."""
    from .transforms import ir
    tr = transforms or ir(100.0, 1e4, 104)
    ao, aux = config2_sizes()
    return synthetic("config2", 6, 16, ao, aux, 1, tr, SEED_BASE + 2,
                     v_scale=coulomb_scale(aux) if v_scale is None else v_scale)


def config3(transforms: Optional[Transforms] = None, n: int = 4, seed: int = SEED_BASE + 3) -> System:
    """Uneven-batching stress system: a Gamma-centred 4x4x4 mesh whose 64
    points are all irreducible (no rotations), each with its own irreps:
    2..10 orbital and 2..10 auxiliary blocks per point, sizes log-uniform in
    [1, 256].  V-block rule (k-dependent blocks): for the pair (k1, k2) with
    momentum transfer q, one block per aux block iq of q, coupling ao block
    ik1 = iq mod n_ao(k1) of k1 to ik2 = (ik1 + iq) mod n_ao(k2) of k2.
    q and -q share their aux blocks.  V at the Coulomb scale 1/sqrt(naux(q))
    of the largest q; beta = 20 on a 6-node Chebyshev grid unless transforms
    are given."""
    from .transforms import chebyshev
    rng = np.random.default_rng(seed)
    pts, kpq, kmq = cubic_mesh(n)
    nk = len(pts)

    def blocks():
        nb = int(rng.integers(2, 11))
        return [int(v) for v in np.clip(np.round(np.exp(rng.uniform(0.0, np.log(256.0), nb))), 1, 256)]
    ao = [blocks() for _ in range(nk)]
    # q and -q share their auxiliary irreps (the Sigma sweep reads a pair with
    # the aux blocks of -q, as P(-q) = P(q)^* relates them)
    aux = [None] * nk
    for q in range(nk):
        mq = int(kmq[0, q])
        if aux[q] is None:
            aux[q] = blocks()
            aux[mq] = list(aux[q])
    V_irreps = {}
    for k1 in range(nk):
        for q in range(nk):
            k2 = int(kpq[k1, q])
            na1, na2 = len(ao[k1]), len(ao[k2])
            V_irreps[k1 * nk + k2] = [(iq, iq % na1, (iq % na1 + iq) % na2) for iq in range(len(aux[q]))]
    tr = transforms or chebyshev(6, 20.0)
    irre = np.arange(nk, dtype=np.int32)
    F = []
    for k in range(nk):
        for b in ao[k]:
            U = haar_unitary(rng, b)
            F.append((U @ np.diag(rng.uniform(-1.0, 1.0, b)) @ U.conj().T).ravel())
    F = np.concatenate(F)[None, :]
    S = np.concatenate([np.eye(b).ravel() for k in range(nk) for b in ao[k]])[None, :].astype(np.complex128)
    return System(nk=nk, ink=nk, ns=1, transforms=tr, ao_sizes=ao, aux_sizes=aux, irre_list=irre, index=irre.copy(),
                  irre_index=irre.copy(), conj_index=irre.copy(), kpq=kpq, kmq=kmq,
                  orep_irreps=[[] for _ in range(nk)], auxrep_irreps=[[] for _ in range(nk)], V_irreps=V_irreps,
                  kweights=np.ones(nk), orep=np.zeros(0, np.complex128), auxrep=np.zeros(0, np.complex128),
                  F=F, S=S, H=np.zeros_like(F), name="config3", v_seed=seed,
                  v_scale=1.0 / float(np.sqrt(max(sum(a) for a in aux))))


def config4_sizes():
    """nao 256 in 6 blocks log-uniform in [16, 80]; naux 1200 in 6 blocks
    log-uniform in [150, 250], redrawn until every block is at most 256 (the
    widest aux block the contraction kernels take)."""
    rng = np.random.default_rng(SEED_BASE + 4)
    ao = split_sizes(rng, 256, 6, 16, 80)
    while True:
        aux = split_sizes(rng, 1200, 6, 150, 250)
        if max(aux) <= 256:
            return ao, aux


def config4(transforms: Optional[Transforms] = None, n: int = 8, ink: int = 29) -> System:
    """8x8x8 mesh reduced to 29 iBZ points (the 35 O_h stars merged as for
    6x6x6); nao 256 in 6 blocks, naux 1200 in 6 blocks; beta = 200 on the IR
    grid of Lambda = 1e5 with 144 points (nts = 146) unless transforms are
    given.  Its V (~5 TB) is streamed (SYMGW_FLAG_STREAM_V)."""
    from .transforms import ir
    tr = transforms or ir(200.0, 1e5, 144, 143)
    ao, aux = config4_sizes()
    return synthetic("config4", n, ink, ao, aux, 1, tr, SEED_BASE + 4, v_scale=coulomb_scale(aux))


def tiny(ns: int = 1, n: int = 2, ink: int = 4, ao=(3, 2), aux=(4, 3), ni: int = 8, beta: float = 10.0,
         seed: int = 7, all_pairs: bool = True) -> System:
    """Small synthetic system for brute-force checks."""
    from .transforms import chebyshev
    return synthetic("tiny", n, ink, list(ao), list(aux), ns, chebyshev(ni, beta), seed, all_pairs=all_pairs)


CONFIG5_NAOS = (26, 52, 104, 156, 256)
CONFIG5_MESHES = (2, 4, 6)


def config5_system(i: int, n: int, nao: int, transforms: Transforms, seed: int = SEED_BASE + 5) -> System:
    """System `i` of the serving mix: a Gamma-centred n^3 mesh reduced to 4 / 8
    / 16 iBZ points (n = 2 / 4 / 6); nao in max(2, min(8, nao // 13)) blocks
    log-uniform in [nao / 2nb, 2 nao / nb]; naux = 4 nao in as many blocks,
    each at most 256; V at the Coulomb scale."""
    rng = np.random.default_rng(seed * 1000 + i)
    nb = max(2, min(8, nao // 13))
    ao = split_sizes(rng, nao, nb, max(1, nao // (2 * nb)), max(4, 2 * nao // nb))
    while True:
        aux = split_sizes(rng, 4 * nao, nb, max(4, 4 * nao // (2 * nb)), min(256, max(8, 8 * nao // nb)))
        if max(aux) <= 256:
            break
    ink = {2: 4, 4: 8, 6: 16}[n]
    return synthetic(f"c5-{i}-{n}^3-nao{nao}", n, ink, ao, aux, 1, transforms, seed * 1000 + i,
                     v_scale=coulomb_scale(aux))


def config5(n_systems: int = 32, seed: int = SEED_BASE + 5, transforms: Optional[Transforms] = None,
            naos=CONFIG5_NAOS, meshes=CONFIG5_MESHES) -> List[System]:
    """The serving mix: `n_systems` systems, nao uniform in {26, 52, 104, 156,
    256} and k-mesh uniform in {2^3, 4^3, 6^3}, each built by config5_system;
    beta = 50 on the IR grid of Lambda = 1e3 with 48 points (nts 50), shared."""
    from .transforms import ir
    tr = transforms or ir(50.0, 1e3, 48)
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_systems):
        n = int(rng.choice(meshes))
        nao = int(rng.choice(naos))
        out.append(config5_system(i, n, nao, tr, seed))
    return out


def serving_mix(n_requests: int = 32, seed: int = SEED_BASE + 5, meshes=(2, 4), naos=(26, 52, 104),
                iterations: int = 2):
    """Seeded request mix for the serving loop: each request is a synthetic
    system on a Gamma-centred n^3 mesh (n from `meshes`, reduced to 4 / 8 / 16
    iBZ points for n = 2 / 4 / 6) with nao from `naos` split into
    max(2, min(8, nao // 13)) blocks and naux = 4 nao in as many blocks;
    beta = 50 on a 30-node Chebyshev grid, ns = 1.  Returns a list of
    (system, iterations)."""
    from .transforms import chebyshev
    rng = np.random.default_rng(seed)
    tr = chebyshev(30, 50.0)
    out = []
    for i in range(n_requests):
        n = int(rng.choice(meshes))
        nao = int(rng.choice(naos))
        nb = max(2, min(8, nao // 13))
        ao = split_sizes(rng, nao, nb, max(2, nao // (2 * nb)), max(4, 2 * nao // nb))
        aux = split_sizes(rng, 4 * nao, nb, max(4, 4 * nao // (2 * nb)), max(8, 8 * nao // nb))
        ink = {2: 4, 4: 8, 6: 16}[n]
        sysm = synthetic(f"mix{i}-n{n}-nao{nao}", n, ink, ao, aux, 1, tr, seed * 100 + i,
                         v_scale=coulomb_scale(aux))
        out.append((sysm, iterations))
    return out
