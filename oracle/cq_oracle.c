/*
 * cq_oracle.c — CPU ORACLE (test infrastructure only).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or the timed CPU baseline. The product
 * (libcqmm.so) never links, loads or calls it.
 *
 * A plain-C restatement of CONQUEST's sparse product for one process, working
 * directly on the reference-layout index arrays of a cqmm_mult_desc:
 *
 *   oracle_mat_mult        mat_mult (src/multiply_module.f90:115-359) with
 *                          loc_trans (:454-539), the B-partition copy of
 *                          Mquest_get_local (src/matrix_comms_module.f90:130-169),
 *                          end_part_comms (:67-103) and the default kernels
 *                          m_kern_max / m_kern_min
 *                          (src/multiply_kernel_default.f90:132-224, 325-412),
 *                          same loop order, same sequential accumulation.
 *   oracle_triplets        the (i,k,j) products the default kernels perform,
 *                          in their loop order (kpart, k, i, j).
 *   oracle_mat_mult_ompgemm  multiply_kernel_ompGemm's m_kern_max / m_kern_min
 *                          (src/multiply_kernel_ompGemm.f90:132-269, 370-494):
 *                          OpenMP schedule(runtime) over i within k, per-(k,i)
 *                          gather -> dgemm -> scatter. This is the timed CPU
 *                          baseline; dgemm comes from the BLAS set with
 *                          oracle_set_blas (numpy's bundled OpenBLAS, ILP64),
 *                          or a plain triple loop if none is set.
 *
 * Parity pins (see tests/): brute-force periodic-image enumeration of the
 * triplet set and C pattern, a dense product of image-resolved blocks on small
 * cells, identity/rank-1 special cases, and agreement between the default and
 * ompGemm restatements.
 */
#include <dlfcn.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/cqmm.h"

/* Fortran 1-based access to a C pointer. */
#define F(arr, i) ((arr)[(i)-1])

typedef void (*dgemm_fn)(const char *, const char *, const int64_t *, const int64_t *, const int64_t *,
                         const double *, const double *, const int64_t *, const double *, const int64_t *,
                         const double *, double *, const int64_t *, size_t, size_t);
typedef void (*setthreads_fn)(int);

static dgemm_fn g_dgemm = NULL;
static char g_blas_name[256] = "naive triple loop";

int oracle_set_blas(const char *path, const char *dgemm_sym, const char *threads_sym) {
  void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  g_dgemm = (dgemm_fn)dlsym(h, dgemm_sym);
  if (!g_dgemm) return -2;
  if (threads_sym) {
    setthreads_fn st = (setthreads_fn)dlsym(h, threads_sym);
    if (st) st(1); /* one BLAS thread per OpenMP thread */
  }
  strncpy(g_blas_name, path, sizeof(g_blas_name) - 1);
  return 0;
}

const char *oracle_blas_name(void) { return g_blas_name; }

/* C(m,n) = alpha*A(m,k)*B(k,n) + beta*C, all 'n','n', column-major. */
static void dgemm_nn(int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                     const double *B, int64_t ldb, double beta, double *Cm, int64_t ldc) {
  if (g_dgemm) {
    g_dgemm("n", "n", &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, Cm, &ldc, 1, 1);
    return;
  }
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double s = 0.0;
      for (int64_t l = 0; l < k; ++l) s += A[i + l * lda] * B[l + j * ldb];
      Cm[i + j * ldc] = alpha * s + beta * Cm[i + j * ldc];
    }
}

/* ---- pieces of mat_mult ------------------------------------------------- */

/* loc_trans (src/multiply_module.f90:454-539); invdir 0: a -> atrans, 1: back. */
static void loc_trans(const cqmm_mult_desc *d, double *a, double *atrans, int invdir) {
  const cqmm_halo *ah = &d->ahalo;
  const cqmm_trans *at = &d->ltrans;
  for (int ni = 1; ni <= ah->ni_in_halo; ++ni) {
    int64_t tpos = F(at->i_nd_beg, ni);
    const int nd2 = F(ah->ndimj, ni);
    for (int nn = 1; nn <= F(at->n_hnab, ni); ++nn) {
      const int i = F(at->i_prim, F(at->i_beg, ni) + nn - 1);
      const int64_t istart = F(ah->i_h2d, (int64_t)(i - 1) * ah->ni_in_halo + ni);
      const int nd1 = F(ah->ndimi, i);
      for (int n1 = 1; n1 <= nd1; ++n1)
        for (int n2 = 1; n2 <= nd2; ++n2) {
          if (invdir == 0)
            F(atrans, (int64_t)(n1 - 1) * nd2 + n2 - 1 + tpos) = F(a, (int64_t)(n2 - 1) * nd1 + n1 - 1 + istart);
          else
            F(a, (int64_t)(n2 - 1) * nd1 + n1 - 1 + istart) = F(atrans, (int64_t)(n1 - 1) * nd2 + n2 - 1 + tpos);
        }
      tpos += (int64_t)nd2 * nd1;
    }
  }
}

/* The B partition of halo partition kpart, as prefetch / Mquest_get_local hand it
 * to the kernel: pointers into bindex and the data offset of the slice. */
typedef struct {
  const int32_t *nbnab, *ibaddr, *ib_nd_acc, *ibseq, *npxyz, *bndim2;
  int64_t b_base; /* b_rem(x) = b(b_base + x) */
  int32_t n_cont, nabs;
} bpart_t;

static void get_bpart(const cqmm_mult_desc *d, int kpart, bpart_t *bp) {
  const cqmm_bindex *B = &d->bmat;
  const int ind_part = F(d->ahalo.lab_hcell, kpart);
  const int ipart = F(d->i_cc2seq, ind_part);
  const int na = F(B->n_atoms, ipart), nabs = F(B->part_nabs, ipart);
  const int64_t s1 = na > 0 ? na : 1, s2 = nabs > 0 ? nabs : 1;
  const int32_t *base = B->bindex + (F(B->array_posn, ipart) - 1);
  bp->nbnab = base;
  bp->ibaddr = base + s1;
  bp->ib_nd_acc = base + 2 * s1;
  bp->ibseq = base + 3 * s1;
  bp->npxyz = base + 3 * s1 + s2;
  bp->bndim2 = base + 3 * s1 + 4 * s2;
  bp->b_base = (int64_t)F(B->nd_offset, ipart) + bp->ib_nd_acc[0] - 1;
  bp->n_cont = na;
  bp->nabs = nabs;
}

/* end_part_comms (src/matrix_comms_module.f90:67-103). */
static int32_t *end_part_comms(const cqmm_mult_desc *d, const bpart_t *bp) {
  const int ncyz = d->ncovery * d->ncoverz;
  int32_t *ibpart = (int32_t *)malloc(sizeof(int32_t) * (bp->nabs > 0 ? bp->nabs : 1));
  int64_t noff = 0;
  for (int k = 1; k <= bp->n_cont; ++k) {
    for (int ni = 1; ni <= F(bp->nbnab, k); ++ni)
      F(ibpart, F(bp->ibaddr, k) + ni - 1) = F(bp->npxyz, noff + 1 + 3 * (ni - 1)) * ncyz +
                                            F(bp->npxyz, noff + 2 + 3 * (ni - 1)) * d->ncoverz +
                                            F(bp->npxyz, noff + 3 + 3 * (ni - 1));
    noff += 3 * (int64_t)F(bp->nbnab, k);
  }
  return ibpart;
}

/* jbnab2ch for atom k of the current partition. */
static void jmap(const cqmm_mult_desc *d, const bpart_t *bp, const int32_t *ibpart, int k_off, int k_in_part,
                 int32_t *jb) {
  const cqmm_halo *ch = &d->chalo;
  const int nbkbeg = F(bp->ibaddr, k_in_part);
  for (int j = 1; j <= F(bp->nbnab, k_in_part); ++j) {
    const int jpart = F(ibpart, nbkbeg + j - 1) + k_off;
    const int jseq = F(bp->ibseq, nbkbeg + j - 1);
    F(jb, j) = F(ch->i_halo, F(ch->i_hbeg, jpart) + jseq - 1);
  }
}

static int max_bnab(const cqmm_mult_desc *d) {
  int mx = 1;
  const cqmm_bindex *B = &d->bmat;
  for (int p = 1; p <= B->n_parts; ++p) {
    const int na = F(B->n_atoms, p);
    const int32_t *base = B->bindex + (F(B->array_posn, p) - 1);
    for (int k = 0; k < na; ++k)
      if (base[k] > mx) mx = base[k];
  }
  return mx;
}

/* ---- default kernels ----------------------------------------------------- */

/* m_kern_max (src/multiply_kernel_default.f90:132-224). */
static void m_kern_max(const cqmm_mult_desc *d, int kpart, const bpart_t *bp, const int32_t *ibpart,
                       const double *a, const double *b, double *c, int32_t *jb) {
  const cqmm_halo *ah = &d->ahalo, *ch = &d->chalo;
  const cqmm_trans *at = &d->ltrans;
  const int k_off = F(ah->lab_hcover, kpart);
  for (int k = 1; k <= F(ah->nh_part, kpart); ++k) {
    const int kh = F(ah->j_beg, kpart) + k - 1;
    const int kin = F(ah->j_seq, kh);
    const int nbkbeg = F(bp->ibaddr, kin);
    const int64_t nb_nd_kbeg = F(bp->ib_nd_acc, kin);
    const int nd3 = F(ah->ndimj, kh);
    jmap(d, bp, ibpart, k_off, kin, jb);
    int64_t nabeg = F(at->i_nd_beg, kh);
    for (int i = 1; i <= F(at->n_hnab, kh); ++i) {
      const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 1);
      const int nd1 = F(ah->ndimi, ip);
      const int64_t icad = (int64_t)(ip - 1) * ch->ni_in_halo;
      int64_t nbbeg = nb_nd_kbeg;
      for (int j = 1; j <= F(bp->nbnab, kin); ++j) {
        const int nd2 = F(bp->bndim2, nbkbeg + j - 1);
        const int jh = F(jb, j);
        if (jh != 0) {
          const int64_t ncbeg = F(ch->i_h2d, icad + jh);
          if (ncbeg != 0) {
            for (int n2 = 1; n2 <= nd2; ++n2) {
              const int64_t nbaddr = bp->b_base + nbbeg + (int64_t)nd3 * (n2 - 1);
              const int64_t ncaddr = ncbeg + (int64_t)nd1 * (n2 - 1);
              for (int n1 = 1; n1 <= nd1; ++n1) {
                const int64_t naaddr = nabeg + (int64_t)nd3 * (n1 - 1);
                for (int n3 = 1; n3 <= nd3; ++n3)
                  F(c, ncaddr + n1 - 1) = F(c, ncaddr + n1 - 1) + F(a, naaddr + n3 - 1) * F(b, nbaddr + n3 - 1);
              }
            }
          }
        }
        nbbeg += (int64_t)nd3 * nd2;
      }
      nabeg += (int64_t)nd1 * nd3;
    }
  }
}

/* m_kern_min (src/multiply_kernel_default.f90:325-412); `a` is the local transpose. */
static void m_kern_min(const cqmm_mult_desc *d, int kpart, const bpart_t *bp, const int32_t *ibpart, double *a,
                       const double *b, const double *c, int32_t *jb) {
  const cqmm_halo *ah = &d->ahalo, *ch = &d->chalo;
  const cqmm_trans *at = &d->ltrans;
  const int k_off = F(ah->lab_hcover, kpart);
  for (int k = 1; k <= F(ah->nh_part, kpart); ++k) {
    const int kh = F(ah->j_beg, kpart) + k - 1;
    const int kin = F(ah->j_seq, kh);
    const int nbkbeg = F(bp->ibaddr, kin);
    const int64_t nb_nd_kbeg = F(bp->ib_nd_acc, kin);
    const int nd3 = F(ah->ndimj, kh);
    jmap(d, bp, ibpart, k_off, kin, jb);
    int64_t nabeg = F(at->i_nd_beg, kh);
    for (int i = 1; i <= F(at->n_hnab, kh); ++i) {
      const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 1);
      const int nd1 = F(ah->ndimi, ip);
      const int64_t icad = (int64_t)(ip - 1) * ch->ni_in_halo;
      int64_t nbbeg = nb_nd_kbeg;
      for (int j = 1; j <= F(bp->nbnab, kin); ++j) {
        const int nd2 = F(bp->bndim2, nbkbeg + j - 1);
        const int jh = F(jb, j);
        if (jh != 0) {
          const int64_t ncbeg = F(ch->i_h2d, icad + jh);
          if (ncbeg != 0) {
            for (int n2 = 1; n2 <= nd2; ++n2) {
              const int64_t nbaddr = bp->b_base + nbbeg + (int64_t)nd3 * (n2 - 1);
              const int64_t ncaddr = ncbeg + (int64_t)nd1 * (n2 - 1);
              for (int n1 = 1; n1 <= nd1; ++n1) {
                const int64_t naaddr = nabeg + (int64_t)nd3 * (n1 - 1);
                for (int n3 = 1; n3 <= nd3; ++n3)
                  F(a, naaddr + n3 - 1) = F(a, naaddr + n3 - 1) + F(c, ncaddr + n1 - 1) * F(b, nbaddr + n3 - 1);
              }
            }
          }
        }
        nbbeg += (int64_t)nd3 * nd2;
      }
      nabeg += (int64_t)nd1 * nd3;
    }
  }
}

/* mat_mult for one process with the default kernels. */
int oracle_mat_mult(const cqmm_mult_desc *d, double *a, const double *b, double *c) {
  double *atrans = (double *)calloc((size_t)d->lena, sizeof(double));
  int32_t *jb = (int32_t *)malloc(sizeof(int32_t) * (size_t)max_bnab(d));
  if (!atrans || !jb) return -1;
  if (d->mult_type == 1) {
    loc_trans(d, a, atrans, 0);
    memset(c, 0, sizeof(double) * (size_t)d->lenc);
  }
  for (int kpart = 1; kpart <= d->ahalo.np_in_halo; ++kpart) {
    bpart_t bp;
    get_bpart(d, kpart, &bp);
    int32_t *ibpart = end_part_comms(d, &bp);
    if (d->mult_type == 1) m_kern_max(d, kpart, &bp, ibpart, atrans, b, c, jb);
    else m_kern_min(d, kpart, &bp, ibpart, atrans, b, c, jb);
    free(ibpart);
  }
  if (d->mult_type == 2) loc_trans(d, a, atrans, 1);
  free(atrans);
  free(jb);
  return 0;
}

/* Products in the default kernels' loop order: (out, p, q) 1-based offsets of
 * the output block, left operand block and right operand block in the arrays
 * mat_mult receives (type 1: c, a, b; type 2: a, c, b). Returns the count;
 * writes at most `cap` entries (pass cap=0 to count). */
int64_t oracle_triplets(const cqmm_mult_desc *d, int64_t *out, int64_t *p, int64_t *q, int64_t cap) {
  const cqmm_halo *ah = &d->ahalo, *ch = &d->chalo;
  const cqmm_trans *at = &d->ltrans;
  int32_t *jb = (int32_t *)malloc(sizeof(int32_t) * (size_t)max_bnab(d));
  int64_t n = 0;
  for (int kpart = 1; kpart <= ah->np_in_halo; ++kpart) {
    bpart_t bp;
    get_bpart(d, kpart, &bp);
    int32_t *ibpart = end_part_comms(d, &bp);
    const int k_off = F(ah->lab_hcover, kpart);
    for (int k = 1; k <= F(ah->nh_part, kpart); ++k) {
      const int kh = F(ah->j_beg, kpart) + k - 1;
      const int kin = F(ah->j_seq, kh);
      const int nbkbeg = F(bp.ibaddr, kin);
      const int nd3 = F(ah->ndimj, kh);
      jmap(d, &bp, ibpart, k_off, kin, jb);
      for (int i = 1; i <= F(at->n_hnab, kh); ++i) {
        const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 1);
        const int64_t icad = (int64_t)(ip - 1) * ch->ni_in_halo;
        const int64_t aoff = F(ah->i_h2d, (int64_t)(ip - 1) * ah->ni_in_halo + kh);
        int64_t nbbeg = F(bp.ib_nd_acc, kin);
        for (int j = 1; j <= F(bp.nbnab, kin); ++j) {
          const int nd2 = F(bp.bndim2, nbkbeg + j - 1);
          const int jh = F(jb, j);
          if (jh != 0) {
            const int64_t ncbeg = F(ch->i_h2d, icad + jh);
            if (ncbeg != 0) {
              if (n < cap) {
                out[n] = d->mult_type == 1 ? ncbeg : aoff;
                p[n] = d->mult_type == 1 ? aoff : ncbeg;
                q[n] = bp.b_base + nbbeg;
              }
              ++n;
            }
          }
          nbbeg += (int64_t)nd3 * nd2;
        }
      }
    }
    free(ibpart);
  }
  free(jb);
  return n;
}

/* ---- multiply_kernel_ompGemm ---------------------------------------------- */

/* m_kern_max, ompGemm (src/multiply_kernel_ompGemm.f90:132-269); called by every
 * thread of the enclosing parallel region. */
static void m_kern_max_gemm(const cqmm_mult_desc *d, int kpart, const bpart_t *bp, const int32_t *ibpart,
                            const double *a, const double *b, double *c, int mx_absb, int mx_halo_nab) {
  const cqmm_halo *ah = &d->ahalo, *ch = &d->chalo;
  const cqmm_trans *at = &d->ltrans;
  const int k_off = F(ah->lab_hcover, kpart);
  int32_t *jb = (int32_t *)malloc(sizeof(int32_t) * mx_absb);
  int64_t *nd1_1st = (int64_t *)malloc(sizeof(int64_t) * (mx_halo_nab + 1));
  int64_t *nd2_1st = (int64_t *)malloc(sizeof(int64_t) * (mx_absb + 1));
  for (int k = 1; k <= F(ah->nh_part, kpart); ++k) {
    const int kh = F(ah->j_beg, kpart) + k - 1;
    const int kin = F(ah->j_seq, kh);
    const int nbkbeg = F(bp->ibaddr, kin);
    const int64_t nb_nd_kbeg = F(bp->ib_nd_acc, kin);
    const int nd3 = F(ah->ndimj, kh);
    const int nbn = F(bp->nbnab, kin);
    F(nd1_1st, 1) = 0;
    for (int i = 2; i <= F(at->n_hnab, kh); ++i) {
      const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 2);
      F(nd1_1st, i) = F(nd1_1st, i - 1) + (int64_t)nd3 * F(ah->ndimi, ip);
    }
    F(nd2_1st, 1) = 0;
    for (int j = 2; j <= nbn; ++j) F(nd2_1st, j) = F(nd2_1st, j - 1) + (int64_t)nd3 * F(bp->bndim2, nbkbeg + j - 2);
    int64_t maxlen = 0;
    jmap(d, bp, ibpart, k_off, kin, jb);
    for (int j = 1; j <= nbn; ++j) maxlen += F(bp->bndim2, nbkbeg + j - 1);
    double *tempb = (double *)malloc(sizeof(double) * (size_t)(nd3 * (maxlen > 0 ? maxlen : 1)));
    double *tempa = (double *)malloc(sizeof(double) * 64 * 64);
    double *tempc = (double *)malloc(sizeof(double) * (size_t)(64 * (maxlen > 0 ? maxlen : 1)));
#pragma omp for schedule(runtime)
    for (int i = 1; i <= F(at->n_hnab, kh); ++i) {
      const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 1);
      const int nd1 = F(ah->ndimi, ip);
      const int64_t nabeg = F(at->i_nd_beg, kh) + F(nd1_1st, i);
      for (int n1 = 1; n1 <= nd1; ++n1)
        for (int n3 = 1; n3 <= nd3; ++n3) tempa[(n1 - 1) + (int64_t)nd1 * (n3 - 1)] = F(a, nabeg + (int64_t)nd3 * (n1 - 1) + n3 - 1);
      const int64_t icad = (int64_t)(ip - 1) * ch->ni_in_halo;
      int64_t sofar = 0;
      for (int j = 1; j <= nbn; ++j) {
        const int nd2 = F(bp->bndim2, nbkbeg + j - 1);
        const int64_t nbbeg = nb_nd_kbeg + F(nd2_1st, j);
        const int jh = F(jb, j);
        if (jh != 0 && F(ch->i_h2d, icad + jh) != 0) {
          for (int n2 = 1; n2 <= nd2; ++n2) {
            const int64_t nbaddr = bp->b_base + nbbeg + (int64_t)nd3 * (n2 - 1);
            for (int n3 = 1; n3 <= nd3; ++n3) tempb[(n3 - 1) + nd3 * (sofar + n2 - 1)] = F(b, nbaddr + n3 - 1);
          }
          sofar += nd2;
        }
      }
      if (sofar > 0) dgemm_nn(nd1, sofar, nd3, 1.0, tempa, nd1, tempb, nd3, 0.0, tempc, nd1);
      sofar = 0;
      for (int j = 1; j <= nbn; ++j) {
        const int nd2 = F(bp->bndim2, nbkbeg + j - 1);
        const int jh = F(jb, j);
        if (jh != 0) {
          const int64_t ncbeg = F(ch->i_h2d, icad + jh);
          if (ncbeg != 0) {
            for (int n2 = 1; n2 <= nd2; ++n2) {
              const int64_t ncaddr = ncbeg + (int64_t)nd1 * (n2 - 1);
              for (int n1 = 1; n1 <= nd1; ++n1) F(c, ncaddr + n1 - 1) += tempc[(n1 - 1) + nd1 * (sofar + n2 - 1)];
            }
            sofar += nd2;
          }
        }
      }
    }
    free(tempb);
    free(tempa);
    free(tempc);
  }
  free(jb);
  free(nd1_1st);
  free(nd2_1st);
}

/* m_kern_min, ompGemm (src/multiply_kernel_ompGemm.f90:370-494). */
static void m_kern_min_gemm(const cqmm_mult_desc *d, int kpart, const bpart_t *bp, const int32_t *ibpart, double *a,
                            const double *b, const double *c, int mx_absb, int mx_halo_nab) {
  const cqmm_halo *ah = &d->ahalo, *ch = &d->chalo;
  const cqmm_trans *at = &d->ltrans;
  const int k_off = F(ah->lab_hcover, kpart);
  int32_t *jb = (int32_t *)malloc(sizeof(int32_t) * mx_absb);
  int64_t *nd1_1st = (int64_t *)malloc(sizeof(int64_t) * (mx_halo_nab + 1));
  int64_t *nd2_1st = (int64_t *)malloc(sizeof(int64_t) * (mx_absb + 1));
  for (int k = 1; k <= F(ah->nh_part, kpart); ++k) {
    const int kh = F(ah->j_beg, kpart) + k - 1;
    const int kin = F(ah->j_seq, kh);
    const int nbkbeg = F(bp->ibaddr, kin);
    const int64_t nb_nd_kbeg = F(bp->ib_nd_acc, kin);
    const int nd3 = F(ah->ndimj, kh);
    const int nbn = F(bp->nbnab, kin);
    F(nd1_1st, 1) = 0;
    for (int i = 2; i <= F(at->n_hnab, kh); ++i) {
      const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 2);
      F(nd1_1st, i) = F(nd1_1st, i - 1) + (int64_t)nd3 * F(ah->ndimi, ip);
    }
    F(nd2_1st, 1) = 0;
    for (int j = 2; j <= nbn; ++j) F(nd2_1st, j) = F(nd2_1st, j - 1) + (int64_t)nd3 * F(bp->bndim2, nbkbeg + j - 2);
    int64_t maxlen = 0;
    jmap(d, bp, ibpart, k_off, kin, jb);
    for (int j = 1; j <= nbn; ++j) maxlen += F(bp->bndim2, nbkbeg + j - 1);
    if (maxlen < 1) maxlen = 1;
#pragma omp for schedule(runtime)
    for (int i = 1; i <= F(at->n_hnab, kh); ++i) {
      const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 1);
      const int nd1 = F(ah->ndimi, ip);
      const int64_t nabeg = F(at->i_nd_beg, kh) + F(nd1_1st, i);
      const int64_t icad = (int64_t)(ip - 1) * ch->ni_in_halo;
      int64_t sofar = 0;
      double *tempb = (double *)malloc(sizeof(double) * (size_t)(nd3 * maxlen));
      double *tempc = (double *)malloc(sizeof(double) * (size_t)(maxlen * nd1));
      for (int j = 1; j <= nbn; ++j) {
        const int nd2 = F(bp->bndim2, nbkbeg + j - 1);
        const int64_t nbbeg = nb_nd_kbeg + F(nd2_1st, j);
        const int jh = F(jb, j);
        if (jh != 0) {
          const int64_t ncbeg = F(ch->i_h2d, icad + jh);
          if (ncbeg != 0) {
            for (int n2 = 1; n2 <= nd2; ++n2) {
              const int64_t nbaddr = bp->b_base + nbbeg + (int64_t)nd3 * (n2 - 1);
              const int64_t ncaddr = ncbeg + (int64_t)nd1 * (n2 - 1);
              for (int n3 = 1; n3 <= nd3; ++n3) tempb[(n3 - 1) + nd3 * (sofar + n2 - 1)] = F(b, nbaddr + n3 - 1);
              for (int n1 = 1; n1 <= nd1; ++n1) tempc[(sofar + n2 - 1) + maxlen * (n1 - 1)] = F(c, ncaddr + n1 - 1);
            }
            sofar += nd2;
          }
        }
      }
      if (sofar > 0) dgemm_nn(nd3, nd1, sofar, 1.0, tempb, nd3, tempc, maxlen, 1.0, &F(a, nabeg), nd3);
      free(tempb);
      free(tempc);
    }
  }
  free(jb);
  free(nd1_1st);
  free(nd2_1st);
}

/* mat_mult with the ompGemm kernels over halo partitions [kp_begin, kp_end)
 * (1-based, kp_end exclusive; pass 1 and np_in_halo+1 for the whole product).
 * A bounded range times a sample of the work; the loc_trans passes always run.
 * loop_seconds (optional) receives the wall time of the partition loop alone. */
int oracle_mat_mult_ompgemm(const cqmm_mult_desc *d, double *a, const double *b, double *c, int kp_begin, int kp_end,
                            double *loop_seconds) {
  double *atrans = (double *)calloc((size_t)d->lena, sizeof(double));
  if (!atrans) return -1;
  const int mx_absb = max_bnab(d);
  int mx_halo_nab = 1;
  for (int kh = 1; kh <= d->ahalo.ni_in_halo; ++kh)
    if (F(d->ltrans.n_hnab, kh) > mx_halo_nab) mx_halo_nab = F(d->ltrans.n_hnab, kh);
  if (d->mult_type == 1) {
    loc_trans(d, a, atrans, 0);
    memset(c, 0, sizeof(double) * (size_t)d->lenc);
  }
  if (kp_begin < 1) kp_begin = 1;
  if (kp_end > d->ahalo.np_in_halo + 1) kp_end = d->ahalo.np_in_halo + 1;
  bpart_t bp;
  int32_t *ibpart = NULL;
  const double t0 = omp_get_wtime();
#pragma omp parallel default(shared)
  {
    for (int kpart = kp_begin; kpart < kp_end; ++kpart) {
#pragma omp master
      {
        get_bpart(d, kpart, &bp);
        if (ibpart) free(ibpart);
        ibpart = end_part_comms(d, &bp);
      }
#pragma omp barrier
      if (d->mult_type == 1) m_kern_max_gemm(d, kpart, &bp, ibpart, atrans, b, c, mx_absb, mx_halo_nab);
      else m_kern_min_gemm(d, kpart, &bp, ibpart, atrans, b, c, mx_absb, mx_halo_nab);
#pragma omp barrier
    }
  }
  if (loop_seconds) *loop_seconds = omp_get_wtime() - t0;
  if (ibpart) free(ibpart);
  if (d->mult_type == 2) loc_trans(d, a, atrans, 1);
  free(atrans);
  return 0;
}

/* Flops of the products in halo partitions [kp_begin, kp_end): sum 2*nd1*nd2*nd3. */
double oracle_flops(const cqmm_mult_desc *d, int kp_begin, int kp_end) {
  const cqmm_halo *ah = &d->ahalo, *ch = &d->chalo;
  const cqmm_trans *at = &d->ltrans;
  int32_t *jb = (int32_t *)malloc(sizeof(int32_t) * (size_t)max_bnab(d));
  double fl = 0.0;
  if (kp_begin < 1) kp_begin = 1;
  if (kp_end > ah->np_in_halo + 1) kp_end = ah->np_in_halo + 1;
  for (int kpart = kp_begin; kpart < kp_end; ++kpart) {
    bpart_t bp;
    get_bpart(d, kpart, &bp);
    int32_t *ibpart = end_part_comms(d, &bp);
    const int k_off = F(ah->lab_hcover, kpart);
    for (int k = 1; k <= F(ah->nh_part, kpart); ++k) {
      const int kh = F(ah->j_beg, kpart) + k - 1;
      const int kin = F(ah->j_seq, kh);
      const int nbkbeg = F(bp.ibaddr, kin);
      const int nd3 = F(ah->ndimj, kh);
      jmap(d, &bp, ibpart, k_off, kin, jb);
      for (int i = 1; i <= F(at->n_hnab, kh); ++i) {
        const int ip = F(at->i_prim, F(at->i_beg, kh) + i - 1);
        const int64_t icad = (int64_t)(ip - 1) * ch->ni_in_halo;
        const int nd1 = F(ah->ndimi, ip);
        for (int j = 1; j <= F(bp.nbnab, kin); ++j) {
          const int jh = F(jb, j);
          if (jh != 0 && F(ch->i_h2d, icad + jh) != 0) fl += 2.0 * nd1 * nd3 * F(bp.bndim2, nbkbeg + j - 1);
        }
      }
    }
    free(ibpart);
  }
  free(jb);
  return fl;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
