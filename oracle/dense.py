"""Dense-tensor restatement of the three contractions for tiny abelian problems -- TEST
INFRASTRUCTURE.  Every symmetry block is scattered into full matrices and the contraction is one
einsum; block structure, schedule tables and coupling bookkeeping play no part, so this pins the
sector-blocked implementations (oracle.cpp and the HIP path) against the plain definitions:

  H psi:   psi'(a',s',r') = sum L_b1(a',a) W_b1b2(s,s') psi(a,s,r) R_b2(r,r')
  left:    L'_b2(r',r)    = sum bra(a',s',r') L_b1(a',a) W_b1b2(s,s') ket(a,s,r)
  right:   R'_b1(a,a')    = sum ket(a,s,r) W_b1b2(s,s') R_b2(r,r') bra(a',s',r')

with W(s,s') the operator entry (row s in the block's left charge, column s' in its right charge)
and transposed (Hermitian-partner) boundary cohorts expanded as conj_scale * partner_block^T
(reference: shtm_tasks / lshtm_tasks / rshtm_tasks and Cohort kernels, solver/tasks.cpp).
"""
import numpy as np


def offsets(index):
    return np.concatenate([[0], np.cumsum(index.sizes)]).astype(int)


def dense_boundary(spec):
    """(aux, Dbra, Dket) array of a BoundarySpec."""
    ob, ok = offsets(spec.bra), offsets(spec.ket)
    out = np.zeros((spec.aux_dim, ob[-1], ok[-1]))
    store = spec.cohort_offsets()
    for c, coh in enumerate(spec.cohorts):
        lb, rb = coh["lr"]
        X, Y = spec.bra.sizes[lb], spec.ket.sizes[rb]
        cs = coh.get("conj_scales", [1.0] * spec.aux_dim)
        for b, o in enumerate(coh["offsets"]):
            if o < 0:
                continue
            if coh.get("tr", 0):
                blk = spec.data[store[c] + o: store[c] + o + X * Y].reshape(X, Y)  # partner block Y x X, col-major
                out[b, ob[lb]:ob[lb] + X, ok[rb]:ok[rb] + Y] = cs[b] * blk
            else:
                blk = spec.data[store[c] + o: store[c] + o + X * Y].reshape(Y, X).T
                out[b, ob[lb]:ob[lb] + X, ok[rb]:ok[rb] + Y] = cs[b] * blk
    return out


def dense_mpo(mpo, phys):
    op_ = offsets(phys)
    d = op_[-1]
    W = np.zeros((mpo.row_dim, mpo.col_dim, d, d))
    for (b1, b2), terms in mpo.terms.items():
        for op, scale in terms:
            for lc, rc, ents in mpo.ops[op].blocks:
                pl, pr = phys.position(lc), phys.position(rc)
                if pl >= len(phys) or pr >= len(phys):
                    continue
                for e in ents:
                    W[b1, b2, op_[pl] + e[0], op_[pr] + e[1]] += scale * e[2]
    return W


def dense_mps(layout, rp):
    ol, op_, orr = offsets(layout.left), offsets(layout.phys), offsets(layout.right)
    psi = np.zeros((ol[-1], op_[-1], orr[-1]))
    for lb, s, rb in layout.segments():
        ls, rs, d = layout.left.sizes[lb], layout.right.sizes[rb], layout.phys.sizes[s]
        blk = layout.rp_block(rp, lb)
        ro = layout.rpb[(layout.phys.charges[s], layout.right.charges[rb])]
        for ss in range(d):
            psi[ol[lb]:ol[lb] + ls, op_[s] + ss, orr[rb]:orr[rb] + rs] = blk[:, ro + ss * rs: ro + (ss + 1) * rs]
    return psi


def from_dense_mps(layout, psi):
    """right-paired vector of the symmetry-allowed entries of a dense (Dl, d, Dr) tensor"""
    ol, op_, orr = offsets(layout.left), offsets(layout.phys), offsets(layout.right)
    rp = np.zeros(layout.n)
    for lb, s, rb in layout.segments():
        ls, rs, d = layout.left.sizes[lb], layout.right.sizes[rb], layout.phys.sizes[s]
        blk = layout.rp_block(rp, lb)
        ro = layout.rpb[(layout.phys.charges[s], layout.right.charges[rb])]
        for ss in range(d):
            blk[:, ro + ss * rs: ro + (ss + 1) * rs] = psi[ol[lb]:ol[lb] + ls, op_[s] + ss, orr[rb]:orr[rb] + rs]
    return rp


def hpsi(mpo, L, R, layout, rp):
    Ld, Rd = dense_boundary(L), dense_boundary(R)
    W = dense_mpo(mpo, layout.phys)
    psi = dense_mps(layout, rp)
    out = np.einsum("xij,xyst,jsk,ykl->itl", Ld, W, psi, Rd, optimize=True)
    return from_dense_mps(layout, out), out


def lstep(mpo, L, bra_layout, bra_rp, ket_layout, ket_rp):
    """(aux, Dbra_right, Dket_right) dense new left boundary"""
    Ld = dense_boundary(L)
    W = dense_mpo(mpo, ket_layout.phys)
    bra, ket = dense_mps(bra_layout, bra_rp), dense_mps(ket_layout, ket_rp)
    return np.einsum("itk,xij,xyst,jsl->ykl", bra, Ld, W, ket, optimize=True)


def rstep(mpo, R, bra_layout, bra_rp, ket_layout, ket_rp):
    """(aux, Dket_left, Dbra_left) dense new right boundary"""
    Rd = dense_boundary(R)
    W = dense_mpo(mpo, ket_layout.phys)
    bra, ket = dense_mps(bra_layout, bra_rp), dense_mps(ket_layout, ket_rp)
    return np.einsum("jsk,xyst,ykl,itl->xji", ket, W, Rd, bra, optimize=True)


def computed_blocks(spec):
    """(b, bra slice, ket slice, block) for every block a step computed (non-transposed cohorts)"""
    ob, ok = offsets(spec.bra), offsets(spec.ket)
    store = spec.cohort_offsets()
    for c, coh in enumerate(spec.cohorts):
        if coh.get("tr", 0):
            continue
        lb, rb = coh["lr"]
        X, Y = spec.bra.sizes[lb], spec.ket.sizes[rb]
        for b, o in enumerate(coh["offsets"]):
            if o < 0:
                continue
            blk = spec.data[store[c] + o: store[c] + o + X * Y].reshape(Y, X).T
            yield b, slice(ob[lb], ob[lb] + X), slice(ok[rb], ok[rb] + Y), blk
