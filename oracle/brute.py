"""Brute-force periodic-image enumeration (oracle; test infrastructure only).

Everything here is computed from atom positions alone, with no cover set, halo
or partition arithmetic, so it independently pins what CONQUEST's indexing
must produce:

  * ``pattern``   — the blocks of a range: (gid_i, gid_j, image shift) with
    |x_j + s*L - x_i|^2 < rcut^2 - RD_ERR, the test of get_naba
    (src/matrix_elements_module.f90:325) with RD_ERR = 1e-8
    (src/numbers_module.f90:96). The image coordinate is formed the way the
    cover set forms it (x + m*dcell with m = s*ngrid, src/cover_module.f90:378-380),
    then differenced, so the comparison is exact.
  * ``triplets``  — the contributing products (i,k,j): type 1 needs A(i,k,s1),
    B(k,j,s2) and C(i,j,s1+s2); type 2 needs X(i,k,s1), B(k,j,s2), Y(i,j,s1+s2)
    (B's row of k evaluated at k's cell copy, as the B partition stores it).
  * ``image_product`` — C(i,j,s) = sum over those triplets of the block products,
    a re-summation in an order unrelated to the reference kernels.
  * ``bloch``     — the k-space form A(q) = sum_s A(s) exp(2 pi i q.s): a full
    product (no truncation) satisfies C(q) = A(q) B(q) exactly.
"""
from __future__ import annotations

import numpy as np

RD_ERR = 1.0e-8


def pattern(xyz, box, ngrid, rcut):
    """Sorted structured array of (gi, gj, sx, sy, sz) for one range."""
    xyz = np.asarray(xyz, np.float64)
    box = np.broadcast_to(np.asarray(box, np.float64), (3,))
    ngrid = np.broadcast_to(np.asarray(ngrid, np.int64), (3,))
    dcell = box / ngrid
    smax = [int(np.ceil(rcut / box[d])) + 1 for d in range(3)]
    rc2 = rcut * rcut - RD_ERR
    out = []
    n = xyz.shape[0]
    gi = np.repeat(np.arange(n), n).reshape(n, n)
    gj = np.tile(np.arange(n), n).reshape(n, n)
    for sx in range(-smax[0], smax[0] + 1):
        for sy in range(-smax[1], smax[1] + 1):
            for sz in range(-smax[2], smax[2] + 1):
                s = (sx, sy, sz)
                img = [xyz[:, d] + float(s[d] * ngrid[d]) * dcell[d] for d in range(3)]
                dx = img[0][None, :] - xyz[:, 0][:, None]
                dy = img[1][None, :] - xyz[:, 1][:, None]
                dz = img[2][None, :] - xyz[:, 2][:, None]
                d2 = dx * dx + dy * dy + dz * dz
                m = d2 < rc2
                if m.any():
                    k = m.sum()
                    out.append(np.stack([gi[m], gj[m], np.full(k, sx), np.full(k, sy), np.full(k, sz)], axis=1))
    arr = np.concatenate(out, axis=0) if out else np.zeros((0, 5), np.int64)
    return arr.astype(np.int64)


def keys(arr, natoms, span=64):
    """Scalar int64 key of (gi, gj, sx, sy, sz) rows (|s| < span/2)."""
    a = np.asarray(arr, np.int64)
    h = span // 2
    return ((((a[:, 0] * natoms + a[:, 1]) * span + (a[:, 2] + h)) * span + (a[:, 3] + h)) * span + (a[:, 4] + h))


def triplets(A, B, Cpat, natoms):
    """Triplet identities as an (n, 8) array: gi, gk, gj, s1(3), ... returned as
    columns (gi, gk, gj, s1x, s1y, s1z, s2x, s2y, s2z) for every (i,k,s1) in A and
    (k,j,s2) in B whose composition (i,j,s1+s2) is in Cpat."""
    ck = np.sort(keys(Cpat, natoms))
    # B rows grouped by k.
    border = np.argsort(B[:, 0], kind="stable")
    Bs = B[border]
    starts = np.searchsorted(Bs[:, 0], np.arange(natoms + 1))
    rows = []
    for gk in range(natoms):
        a = A[A[:, 1] == gk]
        b = Bs[starts[gk]:starts[gk + 1]]
        if len(a) == 0 or len(b) == 0:
            continue
        ia = np.repeat(np.arange(len(a)), len(b))
        ib = np.tile(np.arange(len(b)), len(a))
        comp = np.stack([a[ia, 0], b[ib, 1], a[ia, 2] + b[ib, 2], a[ia, 3] + b[ib, 3], a[ia, 4] + b[ib, 4]], axis=1)
        kk = keys(comp, natoms)
        pos = np.searchsorted(ck, kk)
        pos = np.minimum(pos, len(ck) - 1)
        ok = ck[pos] == kk
        if ok.any():
            rows.append(np.stack([a[ia, 0], np.full(len(ia), gk), b[ib, 1], a[ia, 2], a[ia, 3], a[ia, 4],
                                  b[ib, 2], b[ib, 3], b[ib, 4]], axis=1)[ok])
    return np.concatenate(rows, axis=0) if rows else np.zeros((0, 9), np.int64)


class BlockIndex:
    """Identity -> (offset, nd1, nd2) lookup for one range, from the product's block table."""

    def __init__(self, blocks, natoms):
        ident = np.stack([blocks["gid_i"], blocks["gid_j"], blocks["shift"][:, 0], blocks["shift"][:, 1],
                          blocks["shift"][:, 2]], axis=1).astype(np.int64)
        self.natoms = natoms
        self.k = keys(ident, natoms)
        self.order = np.argsort(self.k)
        self.ks = self.k[self.order]
        self.offset = np.asarray(blocks["offset"], np.int64)
        self.nd1 = np.asarray(blocks["nd1"], np.int64)
        self.nd2 = np.asarray(blocks["nd2"], np.int64)
        self.ident = ident

    def find(self, ident):
        kk = keys(ident, self.natoms)
        pos = np.searchsorted(self.ks, kk)
        pos = np.minimum(pos, len(self.ks) - 1)
        if not np.all(self.ks[pos] == kk):
            raise KeyError("block identity not in pattern")
        return self.order[pos]


def image_product(trip, mult_type, idx_out, idx_p, idx_q, p_vals, q_vals, out_len):
    """Re-sum the products of `trip` (rows from triplets()) blockwise.

    type 1: out C(i,j,s1+s2) = A(i,k,s1) . B(k,j,s2)
    type 2: out X(i,k,s1)    = Y(i,j,s1+s2) . B(k,j,s2)^T
    Returns a new output array of length out_len."""
    out = np.zeros(out_len, np.float64)
    gi, gk, gj = trip[:, 0], trip[:, 1], trip[:, 2]
    s1 = trip[:, 3:6]
    s2 = trip[:, 6:9]
    ik = np.stack([gi, gk, s1[:, 0], s1[:, 1], s1[:, 2]], axis=1)
    kj = np.stack([gk, gj, s2[:, 0], s2[:, 1], s2[:, 2]], axis=1)
    ij = np.stack([gi, gj, s1[:, 0] + s2[:, 0], s1[:, 1] + s2[:, 1], s1[:, 2] + s2[:, 2]], axis=1)
    bq = idx_q.find(kj)
    if mult_type == 1:
        bo, bp = idx_out.find(ij), idx_p.find(ik)
    else:
        bo, bp = idx_out.find(ik), idx_p.find(ij)
    for t in range(len(trip)):
        P = p_vals[idx_p.offset[bp[t]]: idx_p.offset[bp[t]] + idx_p.nd1[bp[t]] * idx_p.nd2[bp[t]]]
        P = P.reshape(idx_p.nd2[bp[t]], idx_p.nd1[bp[t]]).T
        Q = q_vals[idx_q.offset[bq[t]]: idx_q.offset[bq[t]] + idx_q.nd1[bq[t]] * idx_q.nd2[bq[t]]]
        Q = Q.reshape(idx_q.nd2[bq[t]], idx_q.nd1[bq[t]]).T
        R = P @ Q if mult_type == 1 else P @ Q.T
        o = idx_out.offset[bo[t]]
        out[o:o + R.size] += R.T.reshape(-1)
    return out


def bloch(blocks, vals, natoms, nsf_of_gid, qvec):
    """Dense complex matrix sum_s M(s) exp(2 pi i q.s) of one range."""
    off = np.concatenate([[0], np.cumsum(nsf_of_gid)])
    n = int(off[-1])
    M = np.zeros((n, n), np.complex128)
    for b in range(len(blocks["offset"])):
        i, j = blocks["gid_i"][b], blocks["gid_j"][b]
        s = blocks["shift"][b]
        n1, n2 = blocks["nd1"][b], blocks["nd2"][b]
        o = blocks["offset"][b]
        blk = vals[o:o + n1 * n2].reshape(n2, n1).T
        M[off[i]:off[i] + n1, off[j]:off[j] + n2] += blk * np.exp(2j * np.pi * np.dot(qvec, s))
    return M
