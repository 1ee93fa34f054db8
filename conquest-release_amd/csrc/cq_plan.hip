// Device triplet builder: turns CONQUEST's halo / cover-set / B-partition index
// arrays (one process) into an explicit list of contributing (i,k,j) block
// products, grouped by output block in the reference's summation order.
//
// Reference walk reproduced (src/multiply_kernel_default.f90:169-221 for
// m_kern_max, :362-409 for m_kern_min, inside the partition loop of mat_mult,
// src/multiply_module.f90:217-307):
//   for kpart in A-halo partitions, k in kpart:            (halo order)
//     k_in_part = ahalo%j_seq(k); B row of k from the B partition
//       ipart = parts%i_cc2seq(ahalo%lab_hcell(kpart))     (prefetch, :568-623)
//     jbnab2ch(j) = chalo%i_halo(chalo%i_hbeg(ibpart(j)+k_off)+ibseq(j)-1)
//       with ibpart from npxyz (end_part_comms, src/matrix_comms_module.f90:67-103)
//       and k_off = ahalo%lab_hcover(kpart)
//     for i in ltrans%i_prim(k):                            (ascending primary atoms)
//       for j in B row of k:
//         if jbnab2ch(j) /= 0 and chalo%i_h2d(i, jbnab2ch(j)) /= 0: product (i,k,j)
// Type 1 output block is C(i,j) = chalo%i_h2d(i,j); its products are summed in
// k order, so the list is stably sorted by that block. Type 2 output block is
// A(i,k) = ahalo%i_h2d(i,k); its products (j order) are already contiguous.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "../../include/cqmm.h"
#include "cq_common.h"
#include "cq_device.h"
#include "cq_ops.h"
#include "cq_system.h"

namespace cq {

int DevBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 8;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    return hip_check(e, "hipMalloc (plan buffer)");
  }
  bytes = n;
  hbm_track((int64_t)n);
  return CQMM_OK;
}

namespace {
thread_local int64_t t_hbm_bytes = 0, t_hbm_peak = 0;
}

void hbm_track(int64_t delta) {
  t_hbm_bytes += delta;
  if (t_hbm_bytes > t_hbm_peak) t_hbm_peak = t_hbm_bytes;
}
int64_t hbm_current() { return t_hbm_bytes; }
int64_t hbm_peak() { return t_hbm_peak; }
void hbm_peak_reset() { t_hbm_peak = t_hbm_bytes; }

}  // namespace cq


namespace {

using cq::DevBuf;

template <class T>
int upload(DevBuf &d, const T *h, size_t n, hipStream_t s) {
  int st = d.alloc(n * sizeof(T));
  if (st) return st;
  if (n == 0) return CQMM_OK;
  return cq::hip_check(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpy (plan upload)");
}

// Error record written by the builder kernels: first failing code wins.
struct DevErr {
  int code;
  int a, b;
  long long c;
};

enum : int {
  E_NONE = 0,
  E_JPART = 1,     // ibpart + k_off outside the cover set
  E_IHALO = 2,     // i_hbeg(jpart)+jseq-1 outside i_halo
  E_JHALO = 3,     // C-halo number out of range
  E_BRANGE = 4,    // B block past lenb
  E_PRANGE = 5,    // left operand block past its length
  E_BIDX = 6,      // B index slice outside bindex
};

__device__ inline void dev_fail(DevErr *err, int code, int a, int b, long long c) {
  if (atomicCAS(&err->code, 0, code) == 0) {
    err->a = a;
    err->b = b;
    err->c = c;
  }
}

// Internal view of a product's index arrays. The C ABI passes the reference's
// int32 arrays (H = int32_t); the host index mirror passes 64-bit block offsets
// (H = int64_t) for matrices longer than 2^31 elements.
template <class H>
struct HaloV {
  int32_t np_in_halo = 0, ni_in_halo = 0;
  const int32_t *nh_part = nullptr, *j_beg = nullptr, *lab_hcell = nullptr, *lab_hcover = nullptr;
  const int32_t *j_seq = nullptr, *ndimj = nullptr, *i_halo = nullptr, *i_hbeg = nullptr, *ndimi = nullptr;
  const H *i_h2d = nullptr;
  int64_t len_i_halo = 0, len_i_hbeg = 0;
};

template <class H>
struct DescV {
  int32_t mult_type = 0, n_prim = 0, ncovery = 0, ncoverz = 0;
  HaloV<H> ahalo, chalo;
  const int32_t *n_hnab = nullptr, *i_beg = nullptr, *i_prim = nullptr;
  int64_t len_i_prim = 0;
  int32_t b_n_parts = 0;
  const int32_t *b_array_posn = nullptr, *b_n_atoms = nullptr, *b_part_nabs = nullptr, *b_bindex = nullptr;
  const int32_t *b_nd_offset32 = nullptr;
  const int64_t *b_nd_offset64 = nullptr;
  int64_t b_len_bindex = 0;
  const int32_t *i_cc2seq = nullptr;
  int32_t n_cell_parts = 0;
  int64_t lena = 0, lenb = 0, lenc = 0;
  // Symmetric result: transpose pairs of the result's blocks (1-based element
  // offsets, n1 x n2 = dimensions of block p); npairs < 0 for other plans.
  int64_t npairs = -1;
  const int64_t *pair_p = nullptr, *pair_q = nullptr;
  const int32_t *pair_n1 = nullptr, *pair_n2 = nullptr;
  // Tile row order (rank of each primary atom, 0..n_prim-1), or null for
  // primary order.
  const int32_t *row_rank = nullptr;
  int64_t b_nd_offset(int32_t p) const { return b_nd_offset64 ? b_nd_offset64[p] : (int64_t)b_nd_offset32[p]; }
};

// Per halo atom k (host-computed): B-row location and labels.
struct KMeta {
  int32_t k_off;       // ahalo%lab_hcover(kpart)
  int32_t nbnab;       // B neighbours of k
  int32_t nd3;         // ahalo%ndimj(k)
  int32_t kpart;
  int64_t idx_seq;     // 0-based position of k's first i_seq in bindex
  int64_t idx_pxyz;    // 0-based position of k's first npxyz triple
  int64_t idx_ndim;    // 0-based position of k's first ndimj
  int64_t bdata;       // 0-based offset of k's first B block in b
  int64_t jrow;        // start of k's entries in the per-(k,j) arrays
};

// jbnab2ch and the B-block offsets of every (k, j): one thread per halo atom k.
__global__ void k_jmap(const KMeta *__restrict__ km, int32_t ni_a, const int32_t *__restrict__ bindex,
                       int64_t len_bindex, const int32_t *__restrict__ c_i_halo, int64_t len_i_halo,
                       const int32_t *__restrict__ c_i_hbeg, int64_t len_i_hbeg, int32_t ni_c, int32_t ncyz,
                       int32_t ncz, int64_t lenb, int32_t *__restrict__ jh_out, int32_t *__restrict__ nd2_out,
                       int64_t *__restrict__ boff_out, DevErr *err) {
  const int32_t kh = blockIdx.x * blockDim.x + threadIdx.x;
  if (kh >= ni_a) return;
  const KMeta m = km[kh];
  int64_t boff = m.bdata;
  for (int32_t j = 0; j < m.nbnab; ++j) {
    const int64_t is = m.idx_seq + j, ip = m.idx_pxyz + 3 * (int64_t)j, id = m.idx_ndim + j;
    if (is >= len_bindex || ip + 2 >= len_bindex || id >= len_bindex) {
      dev_fail(err, E_BIDX, kh, j, is);
      return;
    }
    const int32_t jseq = bindex[is];
    const int32_t ibpart = bindex[ip] * ncyz + bindex[ip + 1] * ncz + bindex[ip + 2];
    const int32_t jpart = ibpart + m.k_off;
    const int32_t nd2 = bindex[id];
    int32_t jh = 0;
    if (jpart < 1 || jpart > len_i_hbeg) {
      dev_fail(err, E_JPART, kh, j, jpart);
      return;
    }
    const int64_t ih = (int64_t)c_i_hbeg[jpart - 1] + jseq - 1;
    if (ih < 1 || ih > len_i_halo) {
      dev_fail(err, E_IHALO, kh, j, ih);
      return;
    }
    jh = c_i_halo[ih - 1];
    if (jh < 0 || jh > ni_c) {
      dev_fail(err, E_JHALO, kh, j, jh);
      return;
    }
    if (boff + (int64_t)m.nd3 * nd2 > lenb) {
      dev_fail(err, E_BRANGE, kh, j, boff);
      return;
    }
    jh_out[m.jrow + j] = jh;
    nd2_out[m.jrow + j] = nd2;
    boff_out[m.jrow + j] = boff;
    boff += (int64_t)m.nd3 * nd2;
  }
}

// One wavefront per (k, i) item: count contributing j (and the item's flops).
// When `fill` is set, write the products at the item's offset in j order.
template <bool FILL, class H>
__global__ void __launch_bounds__(256) k_items(
    const KMeta *__restrict__ km, const int32_t *__restrict__ item_k, const int32_t *__restrict__ i_prim,
    int64_t n_items, const int32_t *__restrict__ jh_arr, const int32_t *__restrict__ nd2_arr,
    const int64_t *__restrict__ boff_arr, const H *__restrict__ c_i_h2d, int32_t ni_c,
    const H *__restrict__ a_i_h2d, int32_t ni_a, const int32_t *__restrict__ ndimi, int32_t mult_type,
    int64_t len_p, int32_t *__restrict__ counts, int64_t *__restrict__ flops,
    const int64_t *__restrict__ item_off, int32_t *__restrict__ key_out, int64_t *__restrict__ p_out,
    int64_t *__restrict__ q_out, uint8_t *__restrict__ k_out, DevErr *err, const int32_t *__restrict__ crank,
    const int32_t *__restrict__ tmap, int32_t ncg, int64_t nx, uint64_t *__restrict__ tkey,
    uint8_t *__restrict__ tslot, const int32_t *__restrict__ rrank, const uint8_t *__restrict__ omask) {
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (item >= n_items) return;
  const int32_t kh = item_k[item];
  const KMeta m = km[kh];
  const int32_t i = i_prim[item];  // 1-based primary atom
  const int32_t nd1 = ndimi[i - 1];
  const int64_t crow = (int64_t)(i - 1) * ni_c;
  int64_t base = FILL ? item_off[item] : 0;
  int32_t cnt = 0;
  long long fl = 0;
  // Type 1: left operand is A(i,k), the same for every j of the item.
  int64_t pa = 0;
  if (mult_type == 1) pa = (int64_t)a_i_h2d[(int64_t)(i - 1) * ni_a + kh] - 1;
  // Symmetric results skip the mirrored blocks: type 2's result block (i, k)
  // is the item's own; type 1's (i, j) is checked per j.
  const bool item_masked = omask && mult_type == 2 && omask[(int64_t)(i - 1) * ni_a + kh];
  for (int32_t j0 = 0; j0 < m.nbnab; j0 += 64) {
    const int32_t j = j0 + lane;
    bool valid = false;
    int32_t jh = 0, pos = 0, nd2 = 0;
    int64_t ncbeg = 0;
    if (j < m.nbnab) {
      jh = jh_arr[m.jrow + j];
      if (jh != 0) {
        pos = (int32_t)(crow + jh - 1);
        ncbeg = c_i_h2d[crow + jh - 1];
        valid = ncbeg != 0 && !item_masked && !(omask && mult_type == 1 && omask[pos]);
        nd2 = nd2_arr[m.jrow + j];
      }
    }
    if (!FILL && valid && mult_type == 2 && ncbeg - 1 + (int64_t)nd1 * nd2 > len_p)
      dev_fail(err, E_PRANGE, kh, j, ncbeg);
    const unsigned long long mask = __ballot(valid);
    if (FILL && valid) {
      const int32_t before = __popcll(mask & ((1ull << lane) - 1ull));
      const int64_t t = base + cnt + before;
      if (mult_type == 1) {
        key_out[t] = pos;
        p_out[t] = pa;
        if (k_out) k_out[t] = (uint8_t)m.nd3;
      } else {
        p_out[t] = ncbeg - 1;
        if (k_out) k_out[t] = (uint8_t)nd2;
      }
      q_out[t] = boff_arr[m.jrow + j];
      if (tkey) {
        // Tile of the output block (row i, column c) and contraction atom x.
        const int32_t r0 = rrank ? rrank[i - 1] : i - 1;
        const int32_t c = mult_type == 1 ? jh - 1 : kh;
        const int64_t x = mult_type == 1 ? kh : jh - 1;
        const int32_t cr = crank[c];
        const int32_t tile = tmap[(int64_t)(r0 >> 3) * ncg + (cr >> 3)];
        tkey[t] = (uint64_t)tile * (uint64_t)nx + (uint64_t)x;
        tslot[t] = (uint8_t)(((r0 & 7) << 3) | (cr & 7));
      }
    }
    if (!FILL && valid) fl += 2ll * nd1 * nd2 * m.nd3;
    cnt += __popcll(mask);
  }
  if (!FILL) {
    // Wave-reduce the flops.
    for (int off = 32; off > 0; off >>= 1) fl += __shfl_down(fl, off, 64);
    if (lane == 0) {
      counts[item] = cnt;
      flops[item] = fl;
    }
  }
}

// seg_beg[b] = first sorted triplet whose key is >= key of output block b.
__global__ void k_lower_bound(const int32_t *__restrict__ keys, int64_t n, const int32_t *__restrict__ qk,
                              int64_t nq, int64_t *__restrict__ out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nq) return;
  if (b == nq) {
    out[b] = n;
    return;
  }
  const int32_t v = qk[b];
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  out[b] = lo;
}

template <class T>
__global__ void k_gather(const int64_t *__restrict__ src, const int32_t *__restrict__ perm, int64_t n,
                         T *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (T)src[perm[t]];
}

__global__ void k_gather_u8(const uint8_t *__restrict__ src, const int32_t *__restrict__ perm, int64_t n,
                            uint8_t *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[perm[t]];
}

template <class T>
__global__ void k_narrow(const int64_t *__restrict__ src, int64_t n, T *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (T)src[t];
}

__global__ void k_iota(int32_t *__restrict__ v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = (int32_t)t;
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

const char *err_name(int code) {
  switch (code) {
    case E_JPART: return "partition label ibpart+k_off outside the cover set (cover radius below r(A halo)+r(B)?)";
    case E_IHALO: return "i_hbeg(jpart)+jseq-1 outside i_halo";
    case E_JHALO: return "C-halo number out of range";
    case E_BRANGE: return "B block beyond lenb";
    case E_PRANGE: return "left-operand block beyond its length";
    case E_BIDX: return "B index slice outside bindex";
    default: return "unknown";
  }
}

template <class H>
int check_halo(const HaloV<H> &h, int32_t n_prim, const char *name) {
  if (h.np_in_halo <= 0 || h.ni_in_halo <= 0)
    return cq::set_errorf(CQMM_ERR_ARG, "cqmm_plan_create: %s is empty", name);
  if (!h.nh_part || !h.j_beg || !h.lab_hcell || !h.lab_hcover || !h.j_seq || !h.ndimj || !h.i_h2d ||
      !h.i_halo || !h.i_hbeg || !h.ndimi)
    return cq::set_errorf(CQMM_ERR_ARG, "cqmm_plan_create: %s has a null array", name);
  if (h.len_i_halo <= 0 || h.len_i_hbeg <= 0)
    return cq::set_errorf(CQMM_ERR_ARG, "cqmm_plan_create: %s has empty i_halo/i_hbeg", name);
  int64_t acc = 1;
  for (int32_t p = 0; p < h.np_in_halo; ++p) {
    if (h.j_beg[p] != acc || h.nh_part[p] <= 0)
      return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: %s j_beg/nh_part inconsistent at partition %d", name, p + 1);
    acc += h.nh_part[p];
  }
  if (acc - 1 != h.ni_in_halo)
    return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: %s nh_part sums to %lld, ni_in_halo is %d", name,
                          (long long)(acc - 1), h.ni_in_halo);
  for (int32_t p = 0; p < n_prim; ++p)
    if (h.ndimi[p] <= 0 || h.ndimi[p] > 64)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_plan_create: %s ndimi(%d)=%d outside 1..64", name, p + 1, h.ndimi[p]);
  for (int32_t k = 0; k < h.ni_in_halo; ++k)
    if (h.ndimj[k] <= 0 || h.ndimj[k] > 64)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_plan_create: %s ndimj(%d)=%d outside 1..64", name, k + 1, h.ndimj[k]);
  return CQMM_OK;
}

template <class H>
int build_plan(const DescV<H> *d, cqmm_plan **out) {
  *out = nullptr;
  cq::PhaseTimer pt;
  int st = cq::ensure_init();
  if (st) return st;
  hipStream_t s = cq::stream();
  if (d->mult_type != 1 && d->mult_type != 2)
    return cq::set_errorf(CQMM_ERR_ARG, "cqmm_plan_create: mult_type %d (must be 1 or 2)", d->mult_type);
  if (d->n_prim <= 0 || d->ncovery <= 0 || d->ncoverz <= 0 || d->lena <= 0 || d->lenb <= 0 || d->lenc <= 0)
    return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create: non-positive size in descriptor");
  if ((st = check_halo(d->ahalo, d->n_prim, "ahalo"))) return st;
  if ((st = check_halo(d->chalo, d->n_prim, "chalo"))) return st;
  const HaloV<H> &A = d->ahalo, &C = d->chalo;
  struct { const int32_t *n_hnab, *i_beg, *i_prim; int64_t len_i_prim; } T = {d->n_hnab, d->i_beg, d->i_prim, d->len_i_prim};
  struct { int32_t n_parts; const int32_t *array_posn, *n_atoms, *part_nabs, *bindex; int64_t len_bindex; } B = {
      d->b_n_parts, d->b_array_posn, d->b_n_atoms, d->b_part_nabs, d->b_bindex, d->b_len_bindex};
  if (!T.n_hnab || !T.i_beg || !T.i_prim || T.len_i_prim <= 0)
    return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create: ltrans incomplete");
  if (!B.array_posn || !B.n_atoms || !B.part_nabs || (!d->b_nd_offset32 && !d->b_nd_offset64) || !B.bindex ||
      B.n_parts <= 0 || B.len_bindex <= 0)
    return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create: bmat incomplete");
  if (!d->i_cc2seq || d->n_cell_parts <= 0) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create: i_cc2seq missing");
  const int32_t ni_a = A.ni_in_halo, ni_c = C.ni_in_halo;
  if ((int64_t)d->n_prim * ni_a > INT32_MAX || (int64_t)d->n_prim * ni_c > INT32_MAX)
    return cq::set_error(CQMM_ERR_RANGE, "cqmm_plan_create: n_prim*ni_in_halo exceeds int32");

  // ---- per halo atom k: B-row location (prefetch + the k loop header) ----
  std::vector<KMeta> km(ni_a);
  std::vector<int32_t> item_k;
  int64_t jrow = 0, n_items = 0;
  for (int32_t kp = 0; kp < A.np_in_halo; ++kp) {
    const int32_t ind_part = A.lab_hcell[kp];
    if (ind_part < 1 || ind_part > d->n_cell_parts)
      return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: lab_hcell(%d)=%d", kp + 1, ind_part);
    const int32_t ipart = d->i_cc2seq[ind_part - 1];
    if (ipart < 1 || ipart > B.n_parts)
      return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: i_cc2seq(%d)=%d", ind_part, ipart);
    const int32_t na = B.n_atoms[ipart - 1], nabs = B.part_nabs[ipart - 1];
    const int64_t size1 = na > 0 ? na : 1, size2 = nabs > 0 ? nabs : 1;
    const int64_t base = (int64_t)B.array_posn[ipart - 1] - 1;
    if (base < 0 || base + 3 * size1 + 5 * size2 > B.len_bindex)
      return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: B partition %d slice outside bindex", ipart);
    const int32_t first_nd_acc = B.bindex[base + 2 * size1];  // i_nd_acc(1) of the partition
    for (int32_t kk = 0; kk < A.nh_part[kp]; ++kk) {
      const int32_t kh = A.j_beg[kp] - 1 + kk;
      const int32_t kin = A.j_seq[kh];
      if (kin < 1 || kin > na)
        return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: j_seq(%d)=%d outside partition of %d atoms", kh + 1, kin, na);
      KMeta &m = km[kh];
      m.kpart = kp + 1;
      m.k_off = A.lab_hcover[kp];
      m.nbnab = B.bindex[base + (kin - 1)];
      const int32_t ibaddr = B.bindex[base + size1 + (kin - 1)];
      const int32_t ibndacc = B.bindex[base + 2 * size1 + (kin - 1)];
      if (m.nbnab < 0 || ibaddr < 1 || ibaddr - 1 + m.nbnab > size2 || ibndacc < 1)
        return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: B row of halo atom %d inconsistent", kh + 1);
      m.nd3 = A.ndimj[kh];
      m.idx_seq = base + 3 * size1 + (ibaddr - 1);
      m.idx_pxyz = base + 3 * size1 + size2 + 3 * (int64_t)(ibaddr - 1);
      m.idx_ndim = base + 3 * size1 + 4 * size2 + (ibaddr - 1);
      m.bdata = d->b_nd_offset(ipart - 1) + first_nd_acc - 1 + ibndacc - 1;
      m.jrow = jrow;
      jrow += m.nbnab;
    }
  }
  // Items (k, i) in the reference loop order = order of ltrans%i_prim.
  for (int32_t kh = 0; kh < ni_a; ++kh) {
    const int32_t nh = T.n_hnab[kh], ib = T.i_beg[kh];
    if (nh < 0 || ib < 1 || (int64_t)ib - 1 + nh > T.len_i_prim)
      return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: ltrans of halo atom %d outside i_prim", kh + 1);
    if (ib - 1 != n_items)
      return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: ltrans i_beg(%d) not cumulative", kh + 1);
    for (int32_t t = 0; t < nh; ++t) {
      const int32_t i = T.i_prim[ib - 1 + t];
      if (i < 1 || i > d->n_prim) return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: i_prim(%d)=%d", ib + t, i);
      const int64_t aoff = A.i_h2d[(int64_t)(i - 1) * ni_a + kh];
      if (aoff < 1 || aoff - 1 + (int64_t)A.ndimi[i - 1] * A.ndimj[kh] > d->lena)
        return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: A block (%d,%d) offset %lld invalid", i, kh + 1, (long long)aoff);
      item_k.push_back(kh);
    }
    n_items += nh;
  }
  if (n_items == 0) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create: no (k,i) pairs");

  // Uniform support-function count (fast path of the product kernel).
  int uniform = A.ndimi[0];
  for (int32_t p = 0; p < d->n_prim && uniform; ++p)
    if (A.ndimi[p] != uniform || C.ndimi[p] != uniform) uniform = 0;
  for (int32_t k = 0; k < ni_a && uniform; ++k) if (A.ndimj[k] != uniform) uniform = 0;
  for (int32_t k = 0; k < ni_c && uniform; ++k) if (C.ndimj[k] != uniform) uniform = 0;
  // Only these sizes have compile-time kernel instances; others use per-block sizes.
  if (uniform != 4 && uniform != 9 && uniform != 13 && uniform != 16) uniform = 0;
  int maxn = 0;  // largest block dimension (the mixed-size tile kernel takes up to 16)
  for (int32_t p = 0; p < d->n_prim; ++p) maxn = std::max(maxn, std::max(A.ndimi[p], C.ndimi[p]));
  for (int32_t k = 0; k < ni_a; ++k) maxn = std::max(maxn, A.ndimj[k]);
  for (int32_t k = 0; k < ni_c; ++k) maxn = std::max(maxn, C.ndimj[k]);

  // Symmetric result: of each transpose pair (p, q) of result blocks, block q
  // is not computed (its products are dropped) and is written afterwards as
  // the transpose of p. omask marks the result's halo-table positions of the
  // q blocks (type 1: C's table, type 2: A's).
  std::vector<uint8_t> omask;
  std::vector<int64_t> mir_p, mir_q;
  std::vector<uint8_t> mir_n1, mir_n2;
  std::vector<int64_t> skip;  // sorted 0-based offsets of the q blocks
  if (d->npairs >= 0) {
    const int64_t len_out = d->mult_type == 1 ? d->lenc : d->lena;
    if (d->npairs > 0 && (!d->pair_p || !d->pair_q || !d->pair_n1 || !d->pair_n2))
      return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create_symmetric: null pair array");
    for (int64_t t = 0; t < d->npairs; ++t) {
      int64_t pp = d->pair_p[t] - 1, qq = d->pair_q[t] - 1;
      int32_t a = d->pair_n1[t], b = d->pair_n2[t];
      if (a < 1 || b < 1 || a > 64 || b > 64 || pp < 0 || qq < 0 || pp + (int64_t)a * b > len_out ||
          qq + (int64_t)a * b > len_out)
        return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create_symmetric: pair %lld outside the result", (long long)t + 1);
      if (pp == qq) {
        if (a != b) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create_symmetric: non-square self pair");
        continue;
      }
      mir_p.push_back(pp);
      mir_q.push_back(qq);
      mir_n1.push_back((uint8_t)a);
      mir_n2.push_back((uint8_t)b);
      skip.push_back(qq);
    }
    std::sort(skip.begin(), skip.end());
    if (std::adjacent_find(skip.begin(), skip.end()) != skip.end())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create_symmetric: a block is the second block of two pairs");
    std::vector<int64_t> firsts(mir_p);
    std::sort(firsts.begin(), firsts.end());
    // A computed block must not be mirrored by another pair.
    for (int64_t x : firsts)
      if (std::binary_search(skip.begin(), skip.end(), x))
        return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create_symmetric: a block is in two pairs");
    const HaloV<H> &R = d->mult_type == 1 ? C : A;
    const int32_t nr = R.ni_in_halo;
    const int64_t npos = (int64_t)d->n_prim * nr;
    omask.assign((size_t)npos, 0);
    std::atomic<int64_t> found_q(0), found_p(0);
    // Row by row: the row's blocks lie in [lo, hi] of the result, so each
    // lookup searches only that row's part of the sorted offsets.
    cq::parallel_for(d->n_prim, [&](int64_t r0, int64_t r1) {
      int64_t fq = 0, fp = 0;
      for (int64_t row = r0; row < r1; ++row) {
        const H *h2d = R.i_h2d + row * nr;
        int64_t lo = INT64_MAX, hi = -1;
        for (int32_t j = 0; j < nr; ++j)
          if (h2d[j] != 0) {
            lo = std::min<int64_t>(lo, (int64_t)h2d[j] - 1);
            hi = std::max<int64_t>(hi, (int64_t)h2d[j] - 1);
          }
        if (hi < 0) continue;
        const auto sb = std::lower_bound(skip.begin(), skip.end(), lo), se = std::upper_bound(sb, skip.end(), hi);
        const auto fb = std::lower_bound(firsts.begin(), firsts.end(), lo),
                   fe = std::upper_bound(fb, firsts.end(), hi);
        for (int32_t j = 0; j < nr; ++j) {
          const int64_t off = (int64_t)h2d[j] - 1;
          if (off < 0) continue;
          if (std::binary_search(sb, se, off)) {
            omask[row * nr + j] = 1;
            ++fq;
          }
          if (std::binary_search(fb, fe, off)) ++fp;
        }
      }
      found_q += fq;
      found_p += fp;
    }, 1);
    // Every pair must name two blocks of the result, or the mirror would read
    // or write outside them.
    if (found_q.load() != (int64_t)skip.size() || found_p.load() != (int64_t)mir_p.size())
      return cq::set_errorf(CQMM_ERR_ARG,
                            "cqmm_plan_create_symmetric: %lld pairs, of which %lld / %lld blocks are blocks of the result",
                            (long long)mir_p.size(), (long long)found_p.load(), (long long)found_q.load());
  }
  pt.mark("host index checks");
  // ---- device work ----
  DevBuf d_km, d_bindex, d_ci_halo, d_ci_hbeg, d_ci_h2d, d_ai_h2d, d_iprim, d_itemk, d_ndimi, d_err;
  if ((st = upload(d_km, km.data(), km.size(), s))) return st;
  if ((st = upload(d_bindex, B.bindex, (size_t)B.len_bindex, s))) return st;
  if ((st = upload(d_ci_halo, C.i_halo, (size_t)C.len_i_halo, s))) return st;
  if ((st = upload(d_ci_hbeg, C.i_hbeg, (size_t)C.len_i_hbeg, s))) return st;
  if ((st = upload(d_ci_h2d, C.i_h2d, (size_t)d->n_prim * ni_c, s))) return st;
  if ((st = upload(d_ai_h2d, A.i_h2d, (size_t)d->n_prim * ni_a, s))) return st;
  if ((st = upload(d_iprim, T.i_prim, (size_t)n_items, s))) return st;
  if ((st = upload(d_itemk, item_k.data(), item_k.size(), s))) return st;
  if ((st = upload(d_ndimi, A.ndimi, (size_t)d->n_prim, s))) return st;
  DevErr herr = {0, 0, 0, 0};
  if ((st = upload(d_err, &herr, 1, s))) return st;
  DevBuf d_omask;
  if (!omask.empty() && (st = upload(d_omask, omask.data(), omask.size(), s))) return st;

  DevBuf d_jh, d_nd2, d_boff;
  if ((st = d_jh.alloc(std::max<int64_t>(jrow, 1) * 4))) return st;
  if ((st = d_nd2.alloc(std::max<int64_t>(jrow, 1) * 4))) return st;
  if ((st = d_boff.alloc(std::max<int64_t>(jrow, 1) * 8))) return st;
  const int32_t ncyz = d->ncovery * d->ncoverz;
  hipLaunchKernelGGL(k_jmap, dim3(grid1(ni_a, 128)), dim3(128), 0, s, d_km.as<KMeta>(), ni_a,
                     d_bindex.as<int32_t>(), B.len_bindex, d_ci_halo.as<int32_t>(), C.len_i_halo,
                     d_ci_hbeg.as<int32_t>(), C.len_i_hbeg, ni_c, ncyz, d->ncoverz, d->lenb, d_jh.as<int32_t>(),
                     d_nd2.as<int32_t>(), d_boff.as<int64_t>(), d_err.as<DevErr>());
  const int64_t len_p = d->mult_type == 1 ? d->lena : d->lenc;
  DevBuf d_counts, d_flops;
  if ((st = d_counts.alloc(n_items * 4))) return st;
  if ((st = d_flops.alloc(n_items * 8))) return st;
  const unsigned gi = grid1(n_items, 4);
  hipLaunchKernelGGL((k_items<false, H>), dim3(gi), dim3(256), 0, s, d_km.as<KMeta>(), d_itemk.as<int32_t>(),
                     d_iprim.as<int32_t>(), n_items, d_jh.as<int32_t>(), d_nd2.as<int32_t>(), d_boff.as<int64_t>(),
                     d_ci_h2d.as<H>(), ni_c, d_ai_h2d.as<H>(), ni_a, d_ndimi.as<int32_t>(),
                     d->mult_type, len_p, d_counts.as<int32_t>(), d_flops.as<int64_t>(), nullptr, nullptr, nullptr,
                     nullptr, nullptr, d_err.as<DevErr>(), nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr,
                     omask.empty() ? nullptr : d_omask.as<uint8_t>());
  if ((st = cq::hip_check(hipGetLastError(), "builder kernel launch"))) return st;
  std::vector<int32_t> counts(n_items);
  std::vector<int64_t> iflops(n_items);
  if ((st = cq::hip_check(hipMemcpyAsync(counts.data(), d_counts.p, n_items * 4, hipMemcpyDeviceToHost, s), "D2H counts"))) return st;
  if ((st = cq::hip_check(hipMemcpyAsync(iflops.data(), d_flops.p, n_items * 8, hipMemcpyDeviceToHost, s), "D2H flops"))) return st;
  if ((st = cq::hip_check(hipMemcpyAsync(&herr, d_err.p, sizeof(herr), hipMemcpyDeviceToHost, s), "D2H err"))) return st;
  if ((st = cq::hip_check(hipStreamSynchronize(s), "triplet count"))) return st;
  if (herr.code != E_NONE)
    return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: %s (halo atom %d, B neighbour %d, value %lld)",
                          err_name(herr.code), herr.a + 1, herr.b + 1, herr.c);
  pt.mark("upload + triplet count");

  std::vector<int64_t> item_off(n_items + 1, 0);
  double flops = 0.0;
  for (int64_t t = 0; t < n_items; ++t) {
    item_off[t + 1] = item_off[t] + counts[t];
    flops += (double)iflops[t];
  }
  const int64_t n_trip = item_off[n_items];
  std::unique_ptr<cqmm_plan> plan(new cqmm_plan());
  cq::TripletPlan &tp = plan->tp;
  tp.mult_type = d->mult_type;
  tp.q_trans = d->mult_type == 2;
  tp.uniform = uniform;
  tp.maxn = maxn;
  tp.n_trip = n_trip;
  tp.flops = flops;
  tp.len_out = d->mult_type == 1 ? d->lenc : d->lena;
  tp.len_p = len_p;
  tp.len_q = d->lenb;
  tp.off64 = (tp.len_p > INT32_MAX || tp.len_q > INT32_MAX);
  plan->lena = d->lena;
  plan->lenb = d->lenb;
  plan->lenc = d->lenc;
  if (n_trip > INT32_MAX) return cq::set_error(CQMM_ERR_RANGE, "cqmm_plan_create: more than 2^31 block products");

  // ---- output tiles: host part ----
  // Uniform instanced sizes take the fixed-size tile kernel; other sizes up to
  // 16 (mixed species) the mixed-size one.
  bool tiles = (uniform != 0 || (maxn > 0 && maxn <= 16)) && n_trip > 0 && cq::tiles_enabled();
  const int64_t ns2 = (int64_t)uniform * uniform;
  std::vector<int32_t> crank, tmap, rrank;  // rrank: tile row rank of each primary atom (empty: primary order)
  std::vector<int64_t> tout;
  std::vector<uint8_t> trdim, tcdim, xdim;
  int32_t ncg = 0, n_tiles = 0;
  bool class_uniform = false;
  const int64_t nx = d->mult_type == 1 ? ni_a : ni_c;
  // Uniform 9, 13 and 16 may run as one class of the class-tile schedule
  // (CQMM_UNIFORM_TILES=class) instead of the fixed-size tile kernel.
  const bool uclass = uniform != 0 && uniform != 4 && cq::uniform_class_tiles_enabled();
  if (tiles && uniform && !uclass) {
    // Every operand offset must be a whole number of blocks.
    for (int32_t k = 0; k < ni_a && tiles; ++k) tiles = km[k].bdata % ns2 == 0;
    const H *ph2d = d->mult_type == 1 ? A.i_h2d : C.i_h2d;
    const int64_t npn = (int64_t)d->n_prim * (d->mult_type == 1 ? ni_a : ni_c);
    std::atomic<bool> whole(true);
    cq::parallel_for(npn, [&](int64_t b, int64_t e) {
      for (int64_t x = b; x < e; ++x)
        if (ph2d[x] != 0 && (ph2d[x] - 1) % ns2 != 0) {
          whole = false;
          return;
        }
    });
    tiles = whole;
  }
  pt.mark("  tile map: layout checks");
  if (tiles) {
    const HaloV<H> &CH = d->mult_type == 1 ? C : A;  // halo of the output's columns
    crank = cq::halo_morton_rank(CH.np_in_halo, CH.nh_part, CH.lab_hcover, d->ncovery, d->ncoverz);
    const int32_t ncol = CH.ni_in_halo;
    const int32_t *cdims = d->mult_type == 1 ? C.ndimj : A.ndimj;
    int64_t max_rrank = d->n_prim - 1;
    bool class_rows = false;
    if (d->row_rank) {
      // A permutation of the primary atoms (each tile slot holds one block).
      std::vector<uint8_t> seen(d->n_prim, 0);
      for (int32_t p = 0; p < d->n_prim; ++p) {
        const int32_t v = d->row_rank[p];
        if (v < 0 || v >= d->n_prim || seen[v]) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create: bad row order");
        seen[v] = 1;
      }
      rrank.assign(d->row_rank, d->row_rank + d->n_prim);
    }
    if (!uniform && cq::class_tiles_enabled()) {
      class_rows = true;
      // Species mixes: tile rows (primary order) and columns (Morton order)
      // grouped by block size, each group starting at a multiple of 8, so that
      // every tile has one row block size and one column block size.
      std::vector<int32_t> prim(d->n_prim);
      for (int32_t p = 0; p < d->n_prim; ++p) prim[p] = rrank.empty() ? p : rrank[p];
      rrank = cq::group_ranks_by_size(prim, A.ndimi);
      crank = cq::group_ranks_by_size(crank, cdims);
      max_rrank = 0;
      for (int32_t v : rrank) max_rrank = std::max<int64_t>(max_rrank, v);
    }
    int32_t max_crank = 0;
    for (int32_t v : crank) max_crank = std::max(max_crank, v);
    ncg = (max_crank + 8) / 8;
    const int64_t nrg = (max_rrank + 8) / 8;
    auto rowr = [&](int64_t p) -> int64_t { return rrank.empty() ? p : rrank[p]; };
    (void)ncol;
    // Output blocks (row, column, offset) of the product's result.
    std::vector<int64_t> ob_r, ob_c, ob_o;
    if (d->mult_type == 2) {
      // One output block per (k, i) item.
      ob_r.resize(n_items);
      ob_c.resize(n_items);
      ob_o.resize(n_items);
      cq::parallel_for(n_items, [&](int64_t b, int64_t e) {
        for (int64_t t = b; t < e; ++t) {
          const int32_t kh = item_k[t], i = T.i_prim[t];
          ob_r[t] = i - 1;
          ob_c[t] = kh;
          ob_o[t] = (int64_t)A.i_h2d[(int64_t)(i - 1) * ni_a + kh] - 1;
        }
      });
    } else {
      // The blocks of C row by row: count per row, then fill at the row's offset.
      std::vector<int64_t> rbeg((size_t)d->n_prim + 1, 0);
      cq::parallel_for(d->n_prim, [&](int64_t b, int64_t e) {
        for (int64_t p = b; p < e; ++p) {
          int64_t c = 0;
          for (int32_t j = 0; j < ni_c; ++j) c += C.i_h2d[p * ni_c + j] != 0;
          rbeg[p + 1] = c;
        }
      }, 64);
      for (int32_t p = 0; p < d->n_prim; ++p) rbeg[p + 1] += rbeg[p];
      ob_r.resize(rbeg[d->n_prim]);
      ob_c.resize(rbeg[d->n_prim]);
      ob_o.resize(rbeg[d->n_prim]);
      cq::parallel_for(d->n_prim, [&](int64_t b, int64_t e) {
        for (int64_t p = b; p < e; ++p) {
          int64_t k = rbeg[p];
          for (int32_t j = 0; j < ni_c; ++j) {
            const int64_t off = (int64_t)C.i_h2d[p * ni_c + j];
            if (off == 0) continue;
            ob_r[k] = p;
            ob_c[k] = j;
            ob_o[k++] = off - 1;
          }
        }
      }, 64);
    }
    if (!skip.empty()) {
      // Mirrored blocks are no tile's outputs.
      size_t w = 0;
      for (size_t b = 0; b < ob_o.size(); ++b) {
        if (std::binary_search(skip.begin(), skip.end(), ob_o[b])) continue;
        ob_r[w] = ob_r[b];
        ob_c[w] = ob_c[b];
        ob_o[w++] = ob_o[b];
      }
      ob_r.resize(w);
      ob_c.resize(w);
      ob_o.resize(w);
    }
    pt.mark("  tile map: output blocks");
    for (int64_t b = 0; b < (int64_t)ob_o.size() && tiles && uniform; ++b) tiles = ob_o[b] % ns2 == 0;
    if (tiles && nrg * ncg > INT32_MAX) tiles = false;
    if (tiles) {
      tmap.assign((size_t)(nrg * ncg), -1);
      for (size_t b = 0; b < ob_o.size(); ++b) tmap[(rowr(ob_r[b]) >> 3) * ncg + (crank[ob_c[b]] >> 3)] = 0;
      for (auto &v : tmap)
        if (v == 0) v = n_tiles++;
      pt.mark("  tile map: numbering");
      tout.assign((size_t)n_tiles * 64, -1);
      // Block dimensions of each tile row and column, and of each contraction
      // atom (the mixed-size schedule packs blocks by their sizes).
      trdim.assign((size_t)n_tiles * 8, 0);
      tcdim.assign((size_t)n_tiles * 8, 0);
      // Every output block has its own slot: the fill runs in parallel.
      cq::parallel_for((int64_t)ob_o.size(), [&](int64_t b0, int64_t b1) {
        for (int64_t b = b0; b < b1; ++b) {
          const int32_t cr = crank[ob_c[b]];
          const int64_t rr = rowr(ob_r[b]);
          const int32_t tile = tmap[(rr >> 3) * ncg + (cr >> 3)];
          tout[(size_t)tile * 64 + ((rr & 7) << 3) + (cr & 7)] = ob_o[b];
        }
      });
      if (!uniform || uclass) {
        for (size_t b = 0; b < ob_o.size(); ++b) {
          const int32_t cr = crank[ob_c[b]];
          const int64_t rr = rowr(ob_r[b]);
          const int32_t tile = tmap[(rr >> 3) * ncg + (cr >> 3)];
          trdim[(size_t)tile * 8 + (rr & 7)] = (uint8_t)A.ndimi[ob_r[b]];
          tcdim[(size_t)tile * 8 + (cr & 7)] = (uint8_t)cdims[ob_c[b]];
        }
        // One row size and one column size per tile (the class-tile kernel's condition).
        class_uniform = uclass || class_rows;
        for (int32_t t = 0; t < n_tiles && class_uniform; ++t) {
          int rsz = 0, csz = 0;
          for (int k = 0; k < 8; ++k) {
            const int a = trdim[(size_t)t * 8 + k], c = tcdim[(size_t)t * 8 + k];
            if (a && rsz && a != rsz) class_uniform = false;
            if (c && csz && c != csz) class_uniform = false;
            if (a) rsz = a;
            if (c) csz = c;
          }
        }
      }
      const int32_t *xd = d->mult_type == 1 ? A.ndimj : C.ndimj;
      xdim.resize((size_t)nx);
      for (int64_t x = 0; x < nx; ++x) xdim[x] = (uint8_t)xd[x];
    }
  }
  pt.mark("host tile map");
  DevBuf d_crank, d_tmap, d_tkey, d_tslot, d_rrank;
  if (tiles) {
    if ((st = upload(d_crank, crank.data(), crank.size(), s))) return st;
    if (!rrank.empty() && (st = upload(d_rrank, rrank.data(), rrank.size(), s))) return st;
    if ((st = upload(d_tmap, tmap.data(), tmap.size(), s))) return st;
    if ((st = d_tkey.alloc(n_trip * 8))) return st;
    if ((st = d_tslot.alloc(n_trip))) return st;
  }

  DevBuf d_itemoff, d_key, d_p, d_q, d_k;
  if ((st = upload(d_itemoff, item_off.data(), item_off.size(), s))) return st;
  const int64_t nt1 = std::max<int64_t>(n_trip, 1);
  if ((st = d_p.alloc(nt1 * 8))) return st;
  if ((st = d_q.alloc(nt1 * 8))) return st;
  if (d->mult_type == 1 && (st = d_key.alloc(nt1 * 4))) return st;
  if (!uniform && (st = d_k.alloc(nt1))) return st;
  hipLaunchKernelGGL((k_items<true, H>), dim3(gi), dim3(256), 0, s, d_km.as<KMeta>(), d_itemk.as<int32_t>(),
                     d_iprim.as<int32_t>(), n_items, d_jh.as<int32_t>(), d_nd2.as<int32_t>(), d_boff.as<int64_t>(),
                     d_ci_h2d.as<H>(), ni_c, d_ai_h2d.as<H>(), ni_a, d_ndimi.as<int32_t>(),
                     d->mult_type, len_p, nullptr, nullptr, d_itemoff.as<int64_t>(), d_key.as<int32_t>(),
                     d_p.as<int64_t>(), d_q.as<int64_t>(), uniform ? nullptr : d_k.as<uint8_t>(), d_err.as<DevErr>(),
                     tiles ? d_crank.as<int32_t>() : nullptr, tiles ? d_tmap.as<int32_t>() : nullptr, ncg, nx,
                     tiles ? d_tkey.as<uint64_t>() : nullptr, tiles ? d_tslot.as<uint8_t>() : nullptr,
                     tiles && !rrank.empty() ? d_rrank.as<int32_t>() : nullptr,
                     omask.empty() ? nullptr : d_omask.as<uint8_t>());
  if ((st = cq::hip_check(hipGetLastError(), "builder fill launch"))) return st;
  if (tiles) {
    cq::TileBuildInput tin;
    tin.nsf = uclass ? 0 : uniform;
    tin.maxn = maxn;
    tin.tile_rdim = trdim.data();
    tin.tile_cdim = tcdim.data();
    tin.xdim = xdim.data();
    tin.n_trip = n_trip;
    tin.key = d_tkey.as<uint64_t>();
    tin.slot = d_tslot.as<uint8_t>();
    tin.p = d_p.as<int64_t>();
    tin.q = d_q.as<int64_t>();
    tin.nx = nx;
    tin.n_tiles = n_tiles;
    tin.tile_out = tout.data();
    tin.class_uniform = class_uniform;
    pt.mark("triplet fill + tile keys");
    if ((st = cq::build_tiles(tin, tp.tiles, s))) return st;
    pt.mark("tile schedule");
    d_tkey.release();
    d_tslot.release();
    d_tmap.release();
    d_crank.release();
    d_rrank.release();
  }

  const size_t osz = tp.off64 ? 8 : 4;
  if ((st = tp.tp.alloc(nt1 * osz))) return st;
  if ((st = tp.tq.alloc(nt1 * osz))) return st;
  if (d->mult_type == 2) {
    // Segments are the items themselves: every block of the result, once.
    tp.n_seg = n_items;
    std::vector<int64_t> seg_out(n_items);
    std::vector<uint8_t> seg_mn(uniform ? 0 : 2 * n_items);
    cq::parallel_for(n_items, [&](int64_t b, int64_t e) {
      for (int64_t t = b; t < e; ++t) {
        const int32_t kh = item_k[t], i = T.i_prim[t];
        seg_out[t] = (int64_t)A.i_h2d[(int64_t)(i - 1) * ni_a + kh] - 1;
        if (!uniform) {
          seg_mn[2 * t] = (uint8_t)A.ndimi[i - 1];
          seg_mn[2 * t + 1] = (uint8_t)A.ndimj[kh];
        }
      }
    });
    if ((st = upload(tp.seg_beg, item_off.data(), item_off.size(), s))) return st;
    if ((st = upload(tp.seg_out, seg_out.data(), seg_out.size(), s))) return st;
    if (!uniform && (st = upload(tp.seg_mn, seg_mn.data(), seg_mn.size(), s))) return st;
    if (tp.off64) {
      if ((st = cq::hip_check(hipMemcpyAsync(tp.tp.p, d_p.p, n_trip * 8, hipMemcpyDeviceToDevice, s), "copy p"))) return st;
      if ((st = cq::hip_check(hipMemcpyAsync(tp.tq.p, d_q.p, n_trip * 8, hipMemcpyDeviceToDevice, s), "copy q"))) return st;
    } else if (n_trip > 0) {
      hipLaunchKernelGGL(k_narrow<int32_t>, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_p.as<int64_t>(), n_trip, tp.tp.as<int32_t>());
      hipLaunchKernelGGL(k_narrow<int32_t>, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_q.as<int64_t>(), n_trip, tp.tq.as<int32_t>());
    }
    if (!uniform) tp.tk = std::move(d_k);
  } else {
    // Output blocks of C in storage order: nonzero chalo%i_h2d in row-major order.
    // Counted per row of C (checking every offset), then filled at the row's
    // first segment, in parallel.
    const int64_t nc_all = (int64_t)d->n_prim * ni_c;
    std::vector<int64_t> rbeg((size_t)d->n_prim + 1, 0);
    std::atomic<int64_t> bad(INT64_MAX);  // first position with an invalid offset
    cq::parallel_for(d->n_prim, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        int64_t c = 0;
        for (int32_t jh = 0; jh < ni_c; ++jh) {
          const int64_t off = (int64_t)C.i_h2d[i * ni_c + jh];
          if (off == 0) continue;
          ++c;
          if (off < 1 || off - 1 + (int64_t)C.ndimi[i] * C.ndimj[jh] > d->lenc) {
            int64_t cur = bad.load();
            const int64_t pos = i * ni_c + jh;
            while (pos < cur && !bad.compare_exchange_weak(cur, pos)) {
            }
          }
        }
        rbeg[i + 1] = c;
      }
    }, 64);
    if (bad.load() != INT64_MAX) {
      const int64_t pos = bad.load();
      return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: C block (%d,%d) offset %lld invalid",
                            (int)(pos / ni_c) + 1, (int)(pos % ni_c) + 1, (long long)C.i_h2d[pos]);
    }
    for (int32_t i = 0; i < d->n_prim; ++i) rbeg[i + 1] += rbeg[i];
    const int64_t nseg = rbeg[d->n_prim];
    std::vector<int32_t> qkey(nseg);
    std::vector<int64_t> seg_out(nseg);
    std::vector<uint8_t> seg_mn(uniform ? 0 : 2 * nseg);
    cq::parallel_for(d->n_prim, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        int64_t k = rbeg[i];
        for (int32_t jh = 0; jh < ni_c; ++jh) {
          const int64_t off = (int64_t)C.i_h2d[i * ni_c + jh];
          if (off == 0) continue;
          qkey[k] = (int32_t)(i * ni_c + jh);
          seg_out[k] = off - 1;
          if (!uniform) {
            seg_mn[2 * k] = (uint8_t)C.ndimi[i];
            seg_mn[2 * k + 1] = (uint8_t)C.ndimj[jh];
          }
          ++k;
        }
      }
    }, 64);
    tp.n_seg = nseg;
    if ((st = upload(tp.seg_out, seg_out.data(), seg_out.size(), s))) return st;
    if (!uniform && (st = upload(tp.seg_mn, seg_mn.data(), seg_mn.size(), s))) return st;
    // Stable radix sort of the products by output block keeps the k order.
    DevBuf d_idx, d_key2, d_idx2, d_tmp, d_qkey;
    if ((st = d_idx.alloc(nt1 * 4))) return st;
    if ((st = d_key2.alloc(nt1 * 4))) return st;
    if ((st = d_idx2.alloc(nt1 * 4))) return st;
    if (n_trip > 0) hipLaunchKernelGGL(k_iota, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_idx.as<int32_t>(), n_trip);
    unsigned end_bit = 1;
    while (end_bit < 31 && ((int64_t)1 << end_bit) < nc_all) ++end_bit;
    size_t tmp_bytes = 0;
    if (n_trip > 0) {
      hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_key.as<int32_t>(), d_key2.as<int32_t>(),
                                               d_idx.as<int32_t>(), d_idx2.as<int32_t>(), (size_t)n_trip, 0, end_bit, s);
      if ((st = cq::hip_check(e, "radix_sort_pairs (size)"))) return st;
      if ((st = d_tmp.alloc(tmp_bytes))) return st;
      e = rocprim::radix_sort_pairs(d_tmp.p, tmp_bytes, d_key.as<int32_t>(), d_key2.as<int32_t>(), d_idx.as<int32_t>(),
                                    d_idx2.as<int32_t>(), (size_t)n_trip, 0, end_bit, s);
      if ((st = cq::hip_check(e, "radix_sort_pairs"))) return st;
      if (tp.off64) {
        hipLaunchKernelGGL(k_gather<int64_t>, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_p.as<int64_t>(), d_idx2.as<int32_t>(), n_trip, tp.tp.as<int64_t>());
        hipLaunchKernelGGL(k_gather<int64_t>, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_q.as<int64_t>(), d_idx2.as<int32_t>(), n_trip, tp.tq.as<int64_t>());
      } else {
        hipLaunchKernelGGL(k_gather<int32_t>, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_p.as<int64_t>(), d_idx2.as<int32_t>(), n_trip, tp.tp.as<int32_t>());
        hipLaunchKernelGGL(k_gather<int32_t>, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_q.as<int64_t>(), d_idx2.as<int32_t>(), n_trip, tp.tq.as<int32_t>());
      }
      if (!uniform) {
        if ((st = tp.tk.alloc(nt1))) return st;
        hipLaunchKernelGGL(k_gather_u8, dim3(grid1(n_trip, 256)), dim3(256), 0, s, d_k.as<uint8_t>(), d_idx2.as<int32_t>(), n_trip, tp.tk.as<uint8_t>());
      }
    }
    if ((st = upload(d_qkey, qkey.data(), qkey.size(), s))) return st;
    if ((st = tp.seg_beg.alloc((tp.n_seg + 1) * 8))) return st;
    hipLaunchKernelGGL(k_lower_bound, dim3(grid1(tp.n_seg + 1, 256)), dim3(256), 0, s, d_key2.as<int32_t>(), n_trip,
                       d_qkey.as<int32_t>(), tp.n_seg, tp.seg_beg.as<int64_t>());
    if ((st = cq::hip_check(hipGetLastError(), "segment launch"))) return st;
    if ((st = cq::hip_check(hipStreamSynchronize(s), "segment build"))) return st;
  }
  if ((st = cq::hip_check(hipMemcpyAsync(&herr, d_err.p, sizeof(herr), hipMemcpyDeviceToHost, s), "D2H err"))) return st;
  if ((st = cq::hip_check(hipStreamSynchronize(s), "triplet fill"))) return st;
  if (herr.code != E_NONE)
    return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_plan_create: %s (halo atom %d, B neighbour %d, value %lld)",
                          err_name(herr.code), herr.a + 1, herr.b + 1, herr.c);
  pt.mark("segments");
  if (!mir_p.empty()) {
    tp.n_mirror = (int64_t)mir_p.size();
    if ((st = upload(tp.mir_p, mir_p.data(), mir_p.size(), s))) return st;
    if ((st = upload(tp.mir_q, mir_q.data(), mir_q.size(), s))) return st;
    if ((st = upload(tp.mir_n1, mir_n1.data(), mir_n1.size(), s))) return st;
    if ((st = upload(tp.mir_n2, mir_n2.data(), mir_n2.size(), s))) return st;
    if ((st = cq::hip_check(hipStreamSynchronize(s), "mirror pairs"))) return st;
  }
  *out = plan.release();
  return CQMM_OK;
}

template <class H>
void halo_view(const cqmm_halo &h, HaloV<H> &v) {
  v.np_in_halo = h.np_in_halo; v.ni_in_halo = h.ni_in_halo;
  v.nh_part = h.nh_part; v.j_beg = h.j_beg; v.lab_hcell = h.lab_hcell; v.lab_hcover = h.lab_hcover;
  v.j_seq = h.j_seq; v.ndimj = h.ndimj; v.i_halo = h.i_halo; v.i_hbeg = h.i_hbeg; v.ndimi = h.ndimi;
  v.i_h2d = h.i_h2d; v.len_i_halo = h.len_i_halo; v.len_i_hbeg = h.len_i_hbeg;
}

void halo_view64(const cq::RangeIndex &r, HaloV<int64_t> &v) {
  v.np_in_halo = r.np_in_halo; v.ni_in_halo = r.ni_in_halo;
  v.nh_part = r.nh_part.data(); v.j_beg = r.j_beg.data(); v.lab_hcell = r.lab_hcell.data();
  v.lab_hcover = r.lab_hcover.data(); v.j_seq = r.j_seq.data(); v.ndimj = r.ndimj_h.data();
  v.i_halo = r.i_halo.data(); v.i_hbeg = r.i_hbeg.data(); v.ndimi = r.ndimi.data();
  v.i_h2d = r.i_h2d64.data(); v.len_i_halo = (int64_t)r.i_halo.size(); v.len_i_hbeg = (int64_t)r.i_hbeg.size();
}

// Plan from a system's indices; symmetric: with the result range's transpose pairs.
int sys_plan(const cqmm_system *sys, int32_t mult_type, int32_t ra, int32_t rb, int32_t rc, bool symmetric,
             cqmm_plan **out);

}  // namespace

extern "C" {

int cqmm_plan_create(const cqmm_mult_desc *d, cqmm_plan **out) {
  return cqmm_plan_create_symmetric(d, -1, nullptr, nullptr, nullptr, nullptr, out);
}

int cqmm_plan_create_symmetric(const cqmm_mult_desc *d, int64_t npairs, const int64_t *p, const int64_t *q,
                               const int32_t *n1, const int32_t *n2, cqmm_plan **out) {
  return cq::guarded(__func__, [&]() -> int {
    if (!d || !out) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_create: null argument");
    DescV<int32_t> v;
    v.npairs = npairs;
    v.pair_p = p;
    v.pair_q = q;
    v.pair_n1 = n1;
    v.pair_n2 = n2;
    v.mult_type = d->mult_type; v.n_prim = d->n_prim; v.ncovery = d->ncovery; v.ncoverz = d->ncoverz;
    halo_view(d->ahalo, v.ahalo);
    halo_view(d->chalo, v.chalo);
    v.n_hnab = d->ltrans.n_hnab; v.i_beg = d->ltrans.i_beg; v.i_prim = d->ltrans.i_prim; v.len_i_prim = d->ltrans.len_i_prim;
    v.b_n_parts = d->bmat.n_parts; v.b_array_posn = d->bmat.array_posn; v.b_n_atoms = d->bmat.n_atoms;
    v.b_part_nabs = d->bmat.part_nabs; v.b_bindex = d->bmat.bindex; v.b_nd_offset32 = d->bmat.nd_offset;
    v.b_len_bindex = d->bmat.len_bindex;
    v.i_cc2seq = d->i_cc2seq; v.n_cell_parts = d->n_cell_parts;
    v.lena = d->lena; v.lenb = d->lenb; v.lenc = d->lenc;
    return build_plan(&v, out);
  });
}

int cqmm_sys_plan_create(const cqmm_system *sys, int32_t mult_type, int32_t ra, int32_t rb, int32_t rc,
                         cqmm_plan **out) {
  return sys_plan(sys, mult_type, ra, rb, rc, false, out);
}

int cqmm_sys_plan_create_symmetric(const cqmm_system *sys, int32_t mult_type, int32_t ra, int32_t rb, int32_t rc,
                                   cqmm_plan **out) {
  return sys_plan(sys, mult_type, ra, rb, rc, true, out);
}

}  // extern "C"

namespace {

int sys_plan(const cqmm_system *sys, int32_t mult_type, int32_t ra, int32_t rb, int32_t rc, bool symmetric,
             cqmm_plan **out) {
  return cq::guarded("cqmm_sys_plan_create", [&]() -> int {
    if (!sys || !out) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_plan_create: null argument");
    const cq::System &s = sys->s;
    const int32_t nr = (int32_t)s.ranges.size();
    if (ra < 0 || ra >= nr || rb < 0 || rb >= nr || rc < 0 || rc >= nr)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_plan_create: range id out of bounds");
    const cq::RangeIndex &A = s.ranges[ra], &Bm = s.ranges[rb], &Cm = s.ranges[rc];
    if (A.rcut + Bm.rcut > s.rcut_cover + 1e-9)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_plan_create: cover radius %.6g < r(A halo)+r(B) = %.6g",
                            s.rcut_cover, A.rcut + Bm.rcut);
    DescV<int64_t> v;
    v.mult_type = mult_type; v.n_prim = s.natoms; v.ncovery = s.ncover[1]; v.ncoverz = s.ncover[2];
    halo_view64(A, v.ahalo);
    halo_view64(Cm, v.chalo);
    v.n_hnab = A.n_hnab.data(); v.i_beg = A.i_beg.data(); v.i_prim = A.i_prim.data(); v.len_i_prim = (int64_t)A.i_prim.size();
    v.b_n_parts = s.n_parts; v.b_array_posn = Bm.array_posn.data(); v.b_n_atoms = Bm.n_atoms_part.data();
    v.b_part_nabs = Bm.part_nabs.data(); v.b_bindex = Bm.bindex.data(); v.b_nd_offset64 = Bm.nd_offset64.data();
    v.b_len_bindex = (int64_t)Bm.bindex.size();
    v.i_cc2seq = s.i_cc2seq.data(); v.n_cell_parts = s.n_parts;
    v.lena = A.length; v.lenb = Bm.length; v.lenc = Cm.length;
    std::vector<int64_t> pp, qq, blk;
    std::vector<int32_t> a, b, row_rank;
    const int32_t rr = mult_type == 1 ? rc : ra;  // the result's range
    const cq::RangeIndex &O = s.ranges[rr];
    std::vector<int32_t> crank;
    const bool morton = cq::morton_rows_enabled() || (symmetric && !cq::primary_rows_forced());
    if (symmetric || morton) crank = cq::halo_morton_rank(O.np_in_halo, O.nh_part.data(), O.lab_hcover.data(),
                                                          s.ncover[1], s.ncover[2]);
    if (morton) {
      // Tile rows in the order of the result's columns: each primary atom at
      // the column rank (Morton order of the halo, as the tile columns) of its
      // own unshifted copy, so that a tile's rows and columns are the same
      // compact groups of atoms.
      std::vector<int64_t> home(s.natoms, -1);
      for (int32_t h = 0; h < O.ni_in_halo; ++h) {
        const int32_t m = O.halo_cover[h];
        const int32_t *sh = &s.cs_shift[3 * s.c_part[m]];
        if (sh[0] == 0 && sh[1] == 0 && sh[2] == 0) home[s.c_prim[m]] = crank[h];
      }
      if (std::find(home.begin(), home.end(), (int64_t)-1) == home.end()) {
        std::vector<int32_t> order(s.natoms);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return home[x] < home[y]; });
        row_rank.resize(s.natoms);
        for (int32_t k = 0; k < s.natoms; ++k) row_rank[order[k]] = k;
        v.row_rank = row_rank.data();
      }
    }
    if (symmetric) {
      // The transpose pairs of the result's range (C for type 1, A for type 2).
      const int st = cq::transpose_pairs(s, rr, pp, qq, a, b, &blk);
      if (st) return st;
      {
        // Of each pair compute the block in the lower tile (row group, column
        // group): with the rows ordered like the columns, a tile whose rows
        // and columns are different atoms keeps all or none of its blocks.
        auto tile_key = [&](int64_t e) {
          const int32_t row = (int32_t)(std::upper_bound(O.row_beg.begin(), O.row_beg.end(), e) - O.row_beg.begin()) - 1;
          const int32_t h = O.i_halo[O.nb_cover[e]] - 1;
          const int32_t rk = row_rank.empty() ? row : row_rank[row];
          return ((uint64_t)(uint32_t)(rk >> 3) << 32) | (uint32_t)(crank[h] >> 3);
        };
        cq::parallel_for((int64_t)pp.size(), [&](int64_t t0, int64_t t1) {
          for (int64_t t = t0; t < t1; ++t)
            if (tile_key(blk[2 * t + 1]) < tile_key(blk[2 * t])) {
              std::swap(pp[t], qq[t]);
              std::swap(a[t], b[t]);
            }
        });
      }
      v.npairs = (int64_t)pp.size();
      v.pair_p = pp.data();
      v.pair_q = qq.data();
      v.pair_n1 = a.data();
      v.pair_n2 = b.data();
    }
    return build_plan(&v, out);
  });
}

}  // namespace

extern "C" {

int cqmm_plan_destroy(cqmm_plan *plan) {
  return cq::guarded(__func__, [&]() -> int {
    if (plan) {
      (void)hipStreamSynchronize(cq::stream());
      delete plan;
    }
    return CQMM_OK;
  });
}

int cqmm_plan_get_info(const cqmm_plan *plan, cqmm_plan_info *info) {
  return cq::guarded(__func__, [&]() -> int {
    if (!plan || !info) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_get_info: null argument");
    info->n_triplets = plan->tp.n_trip;
    info->n_out_total = plan->tp.n_seg;
    info->flops = plan->tp.flops;
    info->uniform_nsf = plan->tp.uniform;
    info->mult_type = plan->tp.mult_type;
    info->n_tiles = plan->tp.tiles.n_tiles;
    info->n_entries = plan->tp.tiles.n_entries;
    info->n_loads = plan->tp.tiles.n_loads;
    info->tile_family = plan->tp.tiles.n_tiles > 0 ? cq::tile_family(plan->tp.tiles) : 0;
    info->n_mirror = plan->tp.n_mirror;
    info->tile_wgs = plan->tp.tiles.wgs;
    info->kernel = (plan->tp.tiles.n_tiles > 0 && plan->tp.kernel != CQMM_KERNEL_TRIPLET) ? CQMM_KERNEL_TILE
                                                                                            : CQMM_KERNEL_TRIPLET;
    // Output blocks with at least one product.
    std::vector<int64_t> sb(plan->tp.n_seg + 1);
    int st = cq::hip_check(hipMemcpy(sb.data(), plan->tp.seg_beg.p, sb.size() * 8, hipMemcpyDeviceToHost), "D2H seg");
    if (st) return st;
    int64_t nz = 0;
    for (int64_t b = 0; b < plan->tp.n_seg; ++b) nz += sb[b + 1] > sb[b];
    info->n_out_blocks = nz;
    return CQMM_OK;
  });
}

int cqmm_plan_set_kernel(cqmm_plan *plan, int32_t kernel) {
  return cq::guarded(__func__, [&]() -> int {
    if (!plan) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_set_kernel: null plan");
    if (kernel != CQMM_KERNEL_AUTO && kernel != CQMM_KERNEL_TRIPLET && kernel != CQMM_KERNEL_TILE)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_plan_set_kernel: unknown kernel %d", kernel);
    if (kernel == CQMM_KERNEL_TILE && plan->tp.tiles.n_tiles == 0)
      return cq::set_error(CQMM_ERR_STATE, "cqmm_plan_set_kernel: plan has no tile schedule");
    plan->tp.kernel = kernel;
    return CQMM_OK;
  });
}

int cqmm_plan_set_tile_workgroups(cqmm_plan *plan, int32_t wgs) {
  return cq::guarded(__func__, [&]() -> int {
    if (!plan || wgs < 0) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_set_tile_workgroups: bad argument");
    plan->tp.tiles.wgs = wgs;
    return CQMM_OK;
  });
}

int cqmm_plan_tune(cqmm_plan *plan, double *a, int64_t lena, const double *b, int64_t lenb, double *c,
                   int64_t lenc, uint32_t flags, int32_t reps, float *ms_tile, float *ms_triplet) {
  return cq::guarded(__func__, [&]() -> int {
    if (!plan || !ms_tile || !ms_triplet || reps < 1)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_tune: bad argument");
    int st = cq::ensure_init();
    if (st) return st;
    hipStream_t s = cq::stream();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((st = cq::hip_check(hipEventCreate(&e0), "event")) || (st = cq::hip_check(hipEventCreate(&e1), "event"))) {
      if (e0) (void)hipEventDestroy(e0);
      return st;
    }
    auto time_kernel = [&](int32_t kernel, float *ms) -> int {
      plan->tp.kernel = kernel;
      int r = cqmm_mat_mult(plan, a, lena, b, lenb, c, lenc, flags & ~(uint32_t)CQMM_NO_SYNC);
      if (r) return r;
      if ((r = cq::hip_check(hipEventRecord(e0, s), "event record"))) return r;
      for (int32_t k = 0; k < reps; ++k)
        if ((r = cqmm_mat_mult(plan, a, lena, b, lenb, c, lenc, flags | CQMM_NO_SYNC))) return r;
      if ((r = cq::hip_check(hipEventRecord(e1, s), "event record")) ||
          (r = cq::hip_check(hipEventSynchronize(e1), "event sync")) ||
          (r = cq::hip_check(hipEventElapsedTime(ms, e0, e1), "event time")))
        return r;
      *ms /= (float)reps;
      return CQMM_OK;
    };
    *ms_tile = -1.0f;
    *ms_triplet = -1.0f;
    const int32_t before = plan->tp.kernel;
    st = time_kernel(CQMM_KERNEL_TRIPLET, ms_triplet);
    if (!st && plan->tp.tiles.n_tiles > 0) {
      cq::TilePlan &tl = plan->tp.tiles;
      tl.wgs = 0;
      st = time_kernel(CQMM_KERNEL_TILE, ms_tile);
      // Class tiles: persistent grids of 2, 3 and 4 workgroups per CU against
      // one workgroup per tile; the fastest grid stays.
      if (!st && tl.quad) {
        const int32_t n_cu = cq::runtime().n_cu > 0 ? cq::runtime().n_cu : 256;
        int32_t best = 0;
        float best_ms = *ms_tile;
        for (int32_t per_cu = 2; per_cu <= 4 && !st; ++per_cu) {
          if (per_cu * n_cu >= tl.n_tiles) break;
          tl.wgs = per_cu * n_cu;
          float ms = 0.0f;
          st = time_kernel(CQMM_KERNEL_TILE, &ms);
          if (!st && ms < best_ms) {
            best_ms = ms;
            best = tl.wgs;
          }
        }
        tl.wgs = best;
        *ms_tile = best_ms;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) {
      plan->tp.kernel = before;
      return st;
    }
    plan->tp.kernel = (*ms_tile >= 0.0f && *ms_tile <= *ms_triplet) ? CQMM_KERNEL_TILE : CQMM_KERNEL_TRIPLET;
    // The output must hold a result: re-run when the last timed kernel is not the chosen one
    // (all are correct; this keeps the call's result from the kernel and grid that stay).
    if ((plan->tp.kernel == CQMM_KERNEL_TRIPLET && *ms_tile >= 0.0f) || plan->tp.tiles.quad)
      return cqmm_mat_mult(plan, a, lena, b, lenb, c, lenc, flags & ~(uint32_t)CQMM_NO_SYNC);
    return CQMM_OK;
  });
}

int cqmm_plan_get_triplets(const cqmm_plan *plan, int64_t *out, int64_t *p, int64_t *q) {
  return cq::guarded(__func__, [&]() -> int {
    if (!plan || !out || !p || !q) return cq::set_error(CQMM_ERR_ARG, "cqmm_plan_get_triplets: null argument");
    const cq::TripletPlan &tp = plan->tp;
    std::vector<int64_t> sb(tp.n_seg + 1), so(tp.n_seg);
    int st;
    if ((st = cq::hip_check(hipMemcpy(sb.data(), tp.seg_beg.p, sb.size() * 8, hipMemcpyDeviceToHost), "D2H seg"))) return st;
    if ((st = cq::hip_check(hipMemcpy(so.data(), tp.seg_out.p, so.size() * 8, hipMemcpyDeviceToHost), "D2H out"))) return st;
    if (tp.off64) {
      if ((st = cq::hip_check(hipMemcpy(p, tp.tp.p, tp.n_trip * 8, hipMemcpyDeviceToHost), "D2H p"))) return st;
      if ((st = cq::hip_check(hipMemcpy(q, tp.tq.p, tp.n_trip * 8, hipMemcpyDeviceToHost), "D2H q"))) return st;
    } else {
      std::vector<int32_t> pp(tp.n_trip), qq(tp.n_trip);
      if ((st = cq::hip_check(hipMemcpy(pp.data(), tp.tp.p, tp.n_trip * 4, hipMemcpyDeviceToHost), "D2H p"))) return st;
      if ((st = cq::hip_check(hipMemcpy(qq.data(), tp.tq.p, tp.n_trip * 4, hipMemcpyDeviceToHost), "D2H q"))) return st;
      for (int64_t t = 0; t < tp.n_trip; ++t) { p[t] = pp[t]; q[t] = qq[t]; }
    }
    for (int64_t b = 0; b < tp.n_seg; ++b)
      for (int64_t t = sb[b]; t < sb[b + 1]; ++t) out[t] = so[b];
    for (int64_t t = 0; t < tp.n_trip; ++t) { out[t] += 1; p[t] += 1; q[t] += 1; }
    return CQMM_OK;
  });
}

int cqmm_mat_mult(cqmm_plan *plan, double *a, int64_t lena, const double *b, int64_t lenb, double *c,
                  int64_t lenc, uint32_t flags) {
  return cq::guarded(__func__, [&]() -> int {
    if (!plan || !a || !b || !c) return cq::set_error(CQMM_ERR_ARG, "cqmm_mat_mult: null argument");
    if (lena < plan->lena || lenb < plan->lenb || lenc < plan->lenc)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_mat_mult: lengths (%lld,%lld,%lld) below the plan's (%lld,%lld,%lld)",
                            (long long)lena, (long long)lenb, (long long)lenc, (long long)plan->lena,
                            (long long)plan->lenb, (long long)plan->lenc);
    int st = cq::ensure_init();
    if (st) return st;
    hipStream_t s = cq::stream();
    const bool dev = flags & CQMM_DEVICE_PTRS;
    const bool fp32 = flags & CQMM_FP32;
    double *da = a, *dc = c;
    const double *db = b;
    if (!dev) {
      if (plan->sa.bytes < (size_t)plan->lena * 8 && (st = plan->sa.alloc(plan->lena * 8))) return st;
      if (plan->sb.bytes < (size_t)plan->lenb * 8 && (st = plan->sb.alloc(plan->lenb * 8))) return st;
      if (plan->sc.bytes < (size_t)plan->lenc * 8 && (st = plan->sc.alloc(plan->lenc * 8))) return st;
      da = plan->sa.as<double>();
      dc = plan->sc.as<double>();
      db = plan->sb.as<double>();
      if ((st = cq::hip_check(hipMemcpyAsync((void *)db, b, plan->lenb * 8, hipMemcpyHostToDevice, s), "H2D b"))) return st;
      if (plan->tp.mult_type == 1) {
        if ((st = cq::hip_check(hipMemcpyAsync(da, a, plan->lena * 8, hipMemcpyHostToDevice, s), "H2D a"))) return st;
      } else {
        if ((st = cq::hip_check(hipMemcpyAsync(dc, c, plan->lenc * 8, hipMemcpyHostToDevice, s), "H2D c"))) return st;
      }
    }
    // Operands as the kernels see them: P (left), Q (right), O (result).
    const double *dP = plan->tp.mult_type == 1 ? da : dc;
    double *dO = plan->tp.mult_type == 1 ? dc : da;
    if (!fp32) {
      st = cq::launch_block_products(plan->tp, dP, db, dO, false, s);
    } else {
      // FP32 products on FP32 copies; the result is widened back into the FP64 array.
      const cq::TripletPlan &tp = plan->tp;
      if (plan->fp.bytes < (size_t)tp.len_p * 4 && (st = plan->fp.alloc(tp.len_p * 4))) return st;
      if (plan->fq.bytes < (size_t)tp.len_q * 4 && (st = plan->fq.alloc(tp.len_q * 4))) return st;
      if (plan->fo.bytes < (size_t)tp.len_out * 4 && (st = plan->fo.alloc(tp.len_out * 4))) return st;
      const int mode = (flags & CQMM_BF16X3) ? 4 : (flags & CQMM_FP32_RUN64) ? 3 : (flags & CQMM_FP32_ACC64) ? 2 : 1;
      if (mode == 4) {
        if ((st = cq::op_to_bf16x2(tp.len_p, dP, plan->fp.as<uint32_t>(), s))) return st;
        if ((st = cq::op_to_bf16x2(tp.len_q, db, plan->fq.as<uint32_t>(), s))) return st;
      } else {
        if ((st = cq::op_to_f32(tp.len_p, dP, plan->fp.as<float>(), s))) return st;
        if ((st = cq::op_to_f32(tp.len_q, db, plan->fq.as<float>(), s))) return st;
      }
      if ((st = cq::launch_block_products_f32(tp, plan->fp.as<float>(), plan->fq.as<float>(), plan->fo.as<float>(),
                                              mode, s)))
        return st;
      st = cq::op_to_f64(tp.len_out, plan->fo.as<float>(), dO, s);
    }
    if (st) return st;
    if (!dev) {
      if (plan->tp.mult_type == 1) st = cq::hip_check(hipMemcpyAsync(c, dc, plan->lenc * 8, hipMemcpyDeviceToHost, s), "D2H c");
      else st = cq::hip_check(hipMemcpyAsync(a, da, plan->lena * 8, hipMemcpyDeviceToHost, s), "D2H a");
      if (st) return st;
    }
    if (!dev || !(flags & CQMM_NO_SYNC)) return cq::hip_check(hipStreamSynchronize(s), "cqmm_mat_mult");
    return CQMM_OK;
  });
}

}  // extern "C"
