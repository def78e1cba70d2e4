"""Independent checkers used by the tests: brute-force index contractions
(numpy einsum per V block, no GEMM chain) and closed forms."""
import numpy as np


def block_views(sysm):
    """Per full-BZ k: list of (flat offset, size) of ao blocks; same for aux."""
    ao = [list(zip(sysm.ao_block_offsets(k), sysm.ao_sizes[k])) for k in range(sysm.nk)]
    aux = [list(zip(sysm.aux_block_offsets(q), sysm.aux_sizes[q])) for q in range(sysm.nk)]
    return ao, aux


def kao_off_full(sysm):
    o, out = 0, []
    for k in range(sysm.nk):
        out.append(o)
        o += sysm.kslice(k)
    return out


def qaux_off_full(sysm):
    o, out = 0, []
    for q in range(sysm.nk):
        out.append(o)
        o += sysm.qslice(q)
    return out


def brute_P0(sysm, V, G_full, ts):
    """P0[q_ir][t] (q_t_QP layout) by einsum over each V block:
    P0_QP(t) = -(2/ns)/nk sum_{k,s,blocks} G^k_pa(beta-t) V_Qam G^{k+q}_mn(t) conj(V_Ppn)."""
    nk, ns, nts = sysm.nk, sysm.ns, sysm.nts
    ao, aux = block_views(sysm)
    koff = kao_off_full(sysm)
    offs = sysm.kpair_offsets()
    out = sysm.zeros_q_t_QP()
    qoffi = sysm.qaux_off_irre()
    for iqr, q in enumerate(sysm.irre_list):
        q = int(q)
        for t in ts:
            for k in range(nk):
                k2 = int(sysm.kpq[k, q])
                p = k * nk + k2
                if p not in offs:
                    continue
                o = offs[p][0]
                for (iq, i1, i2), (qs, a1, a2) in zip(sysm.V_irreps[p], sysm.V_block_sizes(p)):
                    Vb = V[o:o + qs * a1 * a2].reshape(qs, a1, a2)
                    o += qs * a1 * a2
                    for s in range(ns):
                        g1o, g2o = koff[k] + ao[k][i1][0], koff[k2] + ao[k2][i2][0]
                        G1 = G_full[nts - 1 - t, s, g1o:g1o + a1 * a1].reshape(a1, a1)
                        G2 = G_full[t, s, g2o:g2o + a2 * a2].reshape(a2, a2)
                        blk = np.einsum("pa,Qam,mn,Ppn->QP", G1, Vb, G2, Vb.conj(), optimize=True)
                        base = nts * qoffi[iqr] + t * sysm.qslice(q) + aux[q][iq][0]
                        out[base:base + qs * qs] += blk.ravel()
    out *= -(2.0 / ns) / nk
    return out


def brute_Sigma(sysm, V, G_full, P_full, ts):
    """Sigma[t][s][k_ir] (ts_kij) by einsum over each V block:
    Sigma_ij = -(1/nk) sum_{q,blocks} V_Qim G^{k-q}_mn(t) P^q_QP(t) conj(V_Pjn)."""
    nk, ns, nts = sysm.nk, sysm.ns, sysm.nts
    ao, aux = block_views(sysm)
    koff = kao_off_full(sysm)
    qoff = qaux_off_full(sysm)
    offs = sysm.kpair_offsets()
    out = sysm.zeros_ts_kij()
    koffi = sysm.kao_off_irre()
    for ikr, k in enumerate(sysm.irre_list):
        k = int(k)
        for q in range(nk):
            k2 = int(sysm.kmq[k, q])
            p = k * nk + k2
            if p not in offs:
                continue
            o = offs[p][0]
            for (iq, i1, i2), (qs, a1, a2) in zip(sysm.V_irreps[p], sysm.V_block_sizes(p)):
                Vb = V[o:o + qs * a1 * a2].reshape(qs, a1, a2)
                o += qs * a1 * a2
                for t in ts:
                    pb = nts * qoff[q] + t * sysm.qslice(q) + aux[q][iq][0]
                    Pm = P_full[pb:pb + qs * qs].reshape(qs, qs)
                    for s in range(ns):
                        g2o = koff[k2] + ao[k2][i2][0]
                        G2 = G_full[t, s, g2o:g2o + a2 * a2].reshape(a2, a2)
                        blk = np.einsum("Qim,mn,QP,Pjn->ij", Vb, G2, Pm, Vb.conj(), optimize=True)
                        so = koffi[ikr] + ao[k][i1][0]
                        out[t, s, so:so + a1 * a1] += blk.ravel()
    out *= -1.0 / nk
    return out


def fermi_g(eps, tau, beta):
    """Non-interacting g(tau) = -exp(-eps tau) / (1 + exp(-beta eps)), stable form."""
    eps = np.asarray(eps)[None, :]
    tau = np.asarray(tau)[:, None]
    pos = eps >= 0
    with np.errstate(over="ignore", invalid="ignore"):   # np.where evaluates both branches
        out = np.where(pos, -np.exp(-eps * tau) / (1.0 + np.exp(-beta * np.abs(eps))),
                       -np.exp((beta - tau) * eps) / (1.0 + np.exp(beta * eps)))
    return out


def noninteracting_G_ibz(sysm, mu=0.0):
    """Closed-form G(tau) = U diag(g(eps - mu, tau)) U^H per iBZ block (S = I):
    ts_kij array plus the eigen-decompositions."""
    tau = sysm.transforms.tau
    beta = sysm.transforms.beta
    G = sysm.zeros_ts_kij()
    eig = {}
    koffi = sysm.kao_off_irre()
    for s in range(sysm.ns):
        for ikr, k in enumerate(sysm.irre_list):
            k = int(k)
            for o, b in zip(sysm.ao_block_offsets(k), sysm.ao_sizes[k]):
                F = sysm.F[s, koffi[ikr] + o:koffi[ikr] + o + b * b].reshape(b, b)
                e, U = np.linalg.eigh(F)
                g = fermi_g(e - mu, tau, beta)       # (nts, b)
                Gt = np.einsum("ia,ta,ja->tij", U, g, U.conj())
                G[:, s, koffi[ikr] + o:koffi[ikr] + o + b * b] = Gt.reshape(len(tau), b * b)
                eig[(s, ikr, o)] = (e, U)
    return G, eig


def sigma_hermiticity_error(sysm, Sig):
    """max over (t, s, iBZ k, ao block) of |Sigma_b - Sigma_b^H|, relative to
    max |Sigma|; Sig in ts_kij layout."""
    Sig = np.asarray(Sig).reshape(sysm.nts, sysm.ns, sysm.flat_ao)
    err, offs = 0.0, sysm.kao_off_irre()
    for ik, k in enumerate(sysm.irre_list):
        k = int(k)
        for o, b in zip(sysm.ao_block_offsets(k), sysm.ao_sizes[k]):
            blk = Sig[:, :, offs[ik] + o:offs[ik] + o + b * b].reshape(sysm.nts, sysm.ns, b, b)
            err = max(err, float(np.abs(blk - blk.conj().swapaxes(-1, -2)).max()))
    return err / max(1e-300, float(np.abs(Sig).max()))


def p0_hermiticity_error(sysm, P0):
    """max over (iBZ q, t < nts/2, aux block) of |P0_b - P0_b^H|, relative to
    max |P0|; P0 in q_t_QP layout."""
    P0 = np.asarray(P0)
    err, qoff = 0.0, sysm.qaux_off_irre()
    for iq, q in enumerate(sysm.irre_list):
        q = int(q)
        qs = sysm.qslice(q)
        for t in range(sysm.nts // 2):
            base = sysm.nts * qoff[iq] + t * qs
            for o, b in zip(sysm.aux_block_offsets(q), sysm.aux_sizes[q]):
                blk = P0[base + o:base + o + b * b].reshape(b, b)
                err = max(err, float(np.abs(blk - blk.conj().T).max()))
    return err / max(1e-300, float(np.abs(P0).max()))
