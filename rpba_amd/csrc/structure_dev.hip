// Device-side index structures of a bundle block (see structure_dev.h):
// camera-major order, covisibility pairs grouped by S block, block row and
// column indices, and RobustDelete's compaction.  Sorting is rocPRIM's stable
// LSD radix sort, so the order inside a block is the generation order
// (point, then track positions), the order in which the reference's
// CalcCorrections accumulates its Schur products (rbacore.cpp:1350-1392).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include "structure_dev.h"

#include <cstring>

#include "host_util.h"
#include "kernels.h"

namespace rpba {

#define DCHK(x)                                    \
    do {                                           \
        if ((x) != hipSuccess) return RPBA_E_HIP;  \
    } while (0)

static unsigned bits_for(unsigned long long n)
{
    unsigned b = 1;
    while (b < 63 && (1ull << b) < n) ++b;
    return b;
}

template <class T>
static int grow(T *&p, size_t &cap, size_t n)
{
    if (p && cap >= n) return RPBA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc((void **)&p, sizeof(T) * (n ? n : 1)) != hipSuccess) return RPBA_E_NOMEM;
    cap = n;
    return RPBA_OK;
}

int DevStruct::pinned(size_t n)
{
    if (hpin && hpin_cap >= n) return RPBA_OK;
    if (hpin) (void)hipHostFree(hpin);
    hpin = nullptr;
    hpin_cap = 0;
    if (hipHostMalloc((void **)&hpin, sizeof(int) * (n ? n : 1), hipHostMallocDefault) != hipSuccess) return RPBA_E_NOMEM;
    hpin_cap = n;
    return RPBA_OK;
}

int DevStruct::scratch(size_t bytes)
{
    return grow(tmp, tmp_cap, bytes);
}

// ---------------------------------------------------------------------------

__global__ void k_iota(int n, int *v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// first index j with key[j] >= k*stride (k = 0..nk), sorted keys
__global__ void k_lower_bound(int nk, unsigned stride, const unsigned *keys, int n, int *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nk) return;
    const unsigned long long target = (unsigned long long)k * stride;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((unsigned long long)keys[mid] < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    out[k] = lo;
}

__global__ void k_pair_counts(int npts, const int *toff, int *cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > npts) return;
    if (i == npts) {
        cnt[i] = 0;
        return;
    }
    const int t = toff[i + 1] - toff[i];
    cnt[i] = t * (t - 1) / 2;
}

// one thread per pair: pair p belongs to point i with ppoff[i] <= p <
// ppoff[i+1]; its local index enumerates (x, y), x < y, row by row over the
// point's entries.  key = block (lower camera, higher camera), val = (entry of
// the lower camera << 32) | entry of the higher camera.
__global__ void k_pairs(int npts, int P, unsigned nc, const int *toff, const int *tcam, const int *ppoff, unsigned *key,
                        unsigned long long *val)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int lo = 0, hi = npts;  // last i with ppoff[i] <= p
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ppoff[mid] <= p)
            lo = mid;
        else
            hi = mid;
    }
    const int i = lo;
    const int b = toff[i], t = toff[i + 1] - b;
    int r = p - ppoff[i], x = 0;
    while (r >= t - 1 - x) {
        r -= t - 1 - x;
        ++x;
    }
    const int ex = b + x, ey = b + x + 1 + r;
    const int cx = tcam[ex], cy = tcam[ey];
    const bool sw = cx > cy;
    const unsigned lc = sw ? cy : cx, hc = sw ? cx : cy;
    const unsigned ea = sw ? ey : ex, eb = sw ? ex : ey;
    key[p] = lc * nc + hc;
    val[p] = ((unsigned long long)ea << 32) | eb;
}

__global__ void k_split_blocks(int nb, unsigned nc, const unsigned *ukey, int *bk, int *bl, unsigned *key2, int *bid)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const unsigned k = ukey[b] / nc, l = ukey[b] % nc;
    bk[b] = (int)k;
    bl[b] = (int)l;
    key2[b] = l * nc + k;
    bid[b] = b;
}

__global__ void k_invert(int n, const int *perm, int *inv)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) inv[perm[j]] = j;
}

__global__ void k_set_last(int *a, int idx, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) a[idx] = v;
}

static inline unsigned nb_(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void k_pg4(int ng, const int *pgrp, const int *toff, int4 *pg4)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ng) return;
    const int p0 = pgrp[c], p1 = pgrp[c + 1];
    pg4[c] = make_int4(p0, p1, toff[p0], toff[p1]);
}

int DevStruct::point_groups(hipStream_t st)
{
    npgrp = (ntr + kPtGroup - 1) / kPtGroup;
    if (grow(pgrp, pgrp_cap, (size_t)npgrp + 1)) return RPBA_E_NOMEM;
    if (grow(pg4, pg4_cap, (size_t)std::max(npgrp, 1))) return RPBA_E_NOMEM;
    k_lower_bound<<<nb_(npgrp + 1, 256), 256, 0, st>>>(npgrp, (unsigned)kPtGroup, (const unsigned *)toff, npts, pgrp);
    if (npgrp > 0) k_pg4<<<nb_(npgrp, 256), 256, 0, st>>>(npgrp, pgrp, toff, pg4);
    return RPBA_OK;
}

// chunks of at most kChunk entries of one image over the camera-major list,
// from camoff; reports the first image without entries
int DevStruct::chunks_from_camoff(hipStream_t st, int *first_empty)
{
    int rc = 0;
    *first_empty = -1;
    if (pinned((size_t)nc + 1)) return RPBA_E_NOMEM;
    DCHK(hipMemcpyAsync(hpin, camoff, sizeof(int) * ((size_t)nc + 1), hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    const std::vector<int> co(hpin, hpin + nc + 1);
    for (int k = 0; k < nc; ++k)
        if (co[(size_t)k + 1] == co[k]) {
            *first_empty = k;
            return RPBA_OK;
        }
    std::vector<int> cs, cof((size_t)nc + 1, 0);
    for (int k = 0; k < nc; ++k) {
        cof[k] = (int)cs.size();
        for (int j = co[k]; j < co[(size_t)k + 1]; j += kChunk) cs.push_back(j);
    }
    cof[nc] = (int)cs.size();
    nch = (int)cs.size();
    cs.push_back(ntr);
    rc |= grow(ch_start, ch_start_cap, cs.size());
    rc |= grow(ch_off, ch_off_cap, cof.size());
    rc |= grow(ch4, ch4_cap, (size_t)std::max(nch, 1));
    if (rc) return RPBA_E_NOMEM;
    if (pinned(cs.size() + cof.size() + 4 * (size_t)nch)) return RPBA_E_NOMEM;
    std::copy(cs.begin(), cs.end(), hpin);
    std::copy(cof.begin(), cof.end(), hpin + cs.size());
    int *h4 = hpin + cs.size() + cof.size();
    for (int k = 0; k < nc; ++k)
        for (int c = cof[k]; c < cof[(size_t)k + 1]; ++c) {
            h4[4 * (size_t)c] = cs[c];
            h4[4 * (size_t)c + 1] = cs[(size_t)c + 1] - cs[c];
            h4[4 * (size_t)c + 2] = k;
            h4[4 * (size_t)c + 3] = c == cof[k] ? 1 : 0;
        }
    DCHK(hipMemcpyAsync(ch_start, hpin, sizeof(int) * cs.size(), hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(ch_off, hpin + cs.size(), sizeof(int) * cof.size(), hipMemcpyHostToDevice, st));
    if (nch > 0) DCHK(hipMemcpyAsync(ch4, h4, sizeof(int4) * (size_t)nch, hipMemcpyHostToDevice, st));
    DCHK(hipStreamSynchronize(st));
    return RPBA_OK;
}

int DevStruct::build(hipStream_t st, int *first_empty)
{
    int rc = 0;
    *first_empty = -1;
    if (updated) {
        updated = false;
        *first_empty = updated_empty;
        return RPBA_OK;
    }
    // --- camera-major: stable sort of entry ids by image ---
    rc |= grow(camoff, camoff_cap, (size_t)nc + 1);
    rc |= grow(camq, camq_cap, (size_t)ntr);
    rc |= grow(u32a, u32a_cap, (size_t)ntr);
    rc |= grow(i32a, i32a_cap, (size_t)ntr + 1);
    if (rc) return RPBA_E_NOMEM;
    if (ntr > 0) {
        k_iota<<<nb_(ntr, 256), 256, 0, st>>>(ntr, i32a);
        size_t bytes = 0;
        const unsigned cb = bits_for((unsigned long long)nc);
        DCHK(rocprim::radix_sort_pairs(nullptr, bytes, (const unsigned *)tcam, u32a, i32a, camq, ntr, 0u, cb, st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::radix_sort_pairs(tmp, bytes, (const unsigned *)tcam, u32a, i32a, camq, ntr, 0u, cb, st));
    }
    k_lower_bound<<<nb_(nc + 1, 256), 256, 0, st>>>(nc, 1u, u32a, ntr, camoff);
    if (grow(camrank, camrank_cap, (size_t)std::max(ntr, 1))) return RPBA_E_NOMEM;
    if (ntr > 0) k_invert<<<nb_(ntr, 256), 256, 0, st>>>(ntr, camq, camrank);
    if (point_groups(st)) return RPBA_E_NOMEM;
    rc = chunks_from_camoff(st, first_empty);
    if (rc || *first_empty >= 0) return rc;
    if (!with_pairs) {
        npairs = 0;
        nblk = 0;
        return RPBA_OK;
    }

    // --- covisibility pairs: count, scan, generate, stable sort by block ---
    rc |= grow(ppoff, ppoff_cap, (size_t)npts + 1);
    if (rc) return RPBA_E_NOMEM;
    k_pair_counts<<<nb_(npts + 1, 256), 256, 0, st>>>(npts, toff, i32a);
    {
        size_t bytes = 0;
        DCHK(rocprim::exclusive_scan(nullptr, bytes, i32a, ppoff, 0, (size_t)npts + 1, rocprim::plus<int>(), st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::exclusive_scan(tmp, bytes, i32a, ppoff, 0, (size_t)npts + 1, rocprim::plus<int>(), st));
    }
    DCHK(hipMemcpyAsync(hpin, ppoff + npts, sizeof(int), hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    const int P = hpin[0];
    npairs = P;
    rc |= grow(key_in, key_in_cap, (size_t)P);
    rc |= grow(key_out, key_out_cap, (size_t)P);
    rc |= grow(val_in, val_in_cap, (size_t)P);
    rc |= grow(pab, pab_cap, (size_t)P);
    if (rc) return RPBA_E_NOMEM;
    if (P > 0) k_pairs<<<nb_(P, 256), 256, 0, st>>>(npts, P, (unsigned)nc, toff, tcam, ppoff, key_in, val_in);
    const unsigned kb = bits_for((unsigned long long)nc * (unsigned long long)nc);
    if (P > 0) {
        size_t bytes = 0;
        DCHK(rocprim::radix_sort_pairs(nullptr, bytes, key_in, key_out, val_in, pab, P, 0u, kb, st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::radix_sort_pairs(tmp, bytes, key_in, key_out, val_in, pab, P, 0u, kb, st));
    }

    // --- blocks: run-length encode, offsets, (k, l), row and column indices ---
    rc |= grow(ukey, ukey_cap, (size_t)P + 1);
    rc |= grow(ucnt, ucnt_cap, (size_t)P + 1);
    rc |= grow(nruns, nruns_cap, (size_t)1);
    if (rc) return RPBA_E_NOMEM;
    unsigned nbk = 0;
    if (P > 0) {
        size_t bytes = 0;
        DCHK(rocprim::run_length_encode(nullptr, bytes, key_out, (unsigned)P, ukey, ucnt, nruns, st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::run_length_encode(tmp, bytes, key_out, (unsigned)P, ukey, ucnt, nruns, st));
        DCHK(hipMemcpyAsync(hpin, nruns, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        DCHK(hipStreamSynchronize(st));
        nbk = (unsigned)hpin[0];
    }
    nblk = (int)nbk;
    rc |= grow(poff, poff_cap, (size_t)nblk + 1);
    rc |= grow(bk, bk_cap, (size_t)nblk);
    rc |= grow(bl, bl_cap, (size_t)nblk);
    rc |= grow(roff, roff_cap, (size_t)nc + 1);
    rc |= grow(coff, coff_cap, (size_t)nc + 1);
    rc |= grow(cblk, cblk_cap, (size_t)nblk);
    rc |= grow(key2, key2_cap, (size_t)nblk);
    rc |= grow(key2s, key2s_cap, (size_t)nblk);
    rc |= grow(bid, bid_cap, (size_t)nblk);
    if (rc) return RPBA_E_NOMEM;
    k_set_last<<<1, 1, 0, st>>>((int *)ucnt, nblk, 0);
    {
        size_t bytes = 0;
        DCHK(rocprim::exclusive_scan(nullptr, bytes, (const int *)ucnt, poff, 0, (size_t)nblk + 1, rocprim::plus<int>(), st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::exclusive_scan(tmp, bytes, (const int *)ucnt, poff, 0, (size_t)nblk + 1, rocprim::plus<int>(), st));
    }
    if (nblk > 0) k_split_blocks<<<nb_(nblk, 256), 256, 0, st>>>(nblk, (unsigned)nc, ukey, bk, bl, key2, bid);
    k_lower_bound<<<nb_(nc + 1, 256), 256, 0, st>>>(nc, (unsigned)nc, ukey, nblk, roff);
    if (nblk > 0) {
        size_t bytes = 0;
        DCHK(rocprim::radix_sort_pairs(nullptr, bytes, key2, key2s, bid, cblk, nblk, 0u, kb, st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::radix_sort_pairs(tmp, bytes, key2, key2s, bid, cblk, nblk, 0u, kb, st));
    }
    k_lower_bound<<<nb_(nc + 1, 256), 256, 0, st>>>(nc, (unsigned)nc, key2s, nblk, coff);
    rc = strip_slabs(st);
    if (rc) return rc;
    DCHK(hipStreamSynchronize(st));
    return RPBA_OK;
}

__global__ void k_strip_keys(int nb, unsigned nc, const int *bk, const int *bl, unsigned *key, int *bid)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    key[b] = (unsigned)(bk[b] / kSlab) * nc + (unsigned)bl[b];
    bid[b] = b;
}

// per slab: column, rows present, and the slab of each of its slots
__global__ void k_slab_info(int ns, unsigned nc, const unsigned *skey, const int *bs, const int *sperm, const int *bk,
                            int *sl, int *smask, int *slot_slab)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    sl[s] = (int)(skey[s] % nc);
    int m = 0;
    for (int q = bs[s]; q < bs[s + 1]; ++q) {
        m |= 1 << (bk[sperm[q]] % kSlab);
        slot_slab[q] = s;
    }
    smask[s] = m;
}

__global__ void k_chunk_counts(int nstrip, int T, const int *strip_s0, int *cnt)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nstrip) return;
    cnt[s] = (s == nstrip) ? 0 : (strip_s0[s + 1] - strip_s0[s] + T - 1) / T;
}

__global__ void k_chunk_fill(int nstrip, const int *chunk_off, int *chunk_strip)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstrip) return;
    for (int c = chunk_off[s]; c < chunk_off[s + 1]; ++c) chunk_strip[c] = s;
}

// Strip slabs of the upper blocks (after the row-major block list exists).
int DevStruct::strip_slabs(hipStream_t st)
{
    int rc = 0;
    nstrip = (nc + kSlab - 1) / kSlab;
    const size_t nb1 = (size_t)std::max(nblk, 1);
    rc |= grow(sslot, sslot_cap, nb1);
    rc |= grow(sperm, sperm_cap, nb1);
    rc |= grow(slot_slab, slot_slab_cap, nb1);
    rc |= grow(slab_key, slab_key_cap, nb1);
    rc |= grow(slab_cnt, slab_cnt_cap, nb1 + 1);
    rc |= grow(strip_s0, strip_s0_cap, (size_t)nstrip + 1);
    rc |= grow(chunk_off, chunk_off_cap, (size_t)nstrip + 1);
    rc |= grow(cs_off, cs_off_cap, (size_t)nc + 1);
    rc |= grow(i32a, i32a_cap, (size_t)nstrip + 1);
    if (rc) return RPBA_E_NOMEM;
    nslab = 0;
    if (nblk > 0) {
        // stable sort of the row-major block ids by (strip, column): rows stay ascending inside a slab
        k_strip_keys<<<nb_(nblk, 256), 256, 0, st>>>(nblk, (unsigned)nc, bk, bl, key2, bid);
        const unsigned kb = bits_for((unsigned long long)nstrip * (unsigned long long)nc);
        size_t bytes = 0;
        DCHK(rocprim::radix_sort_pairs(nullptr, bytes, key2, key2s, bid, sperm, nblk, 0u, kb, st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::radix_sort_pairs(tmp, bytes, key2, key2s, bid, sperm, nblk, 0u, kb, st));
        bytes = 0;
        DCHK(rocprim::run_length_encode(nullptr, bytes, key2s, (unsigned)nblk, slab_key, slab_cnt, nruns, st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::run_length_encode(tmp, bytes, key2s, (unsigned)nblk, slab_key, slab_cnt, nruns, st));
        DCHK(hipMemcpyAsync(hpin, nruns, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        DCHK(hipStreamSynchronize(st));
        nslab = (int)((unsigned)hpin[0]);
    }
    const size_t ns1 = (size_t)std::max(nslab, 1);
    rc |= grow(slab_bs, slab_bs_cap, ns1 + 1);
    rc |= grow(slab_l, slab_l_cap, ns1);
    rc |= grow(slab_mask, slab_mask_cap, ns1);
    rc |= grow(cs_list, cs_list_cap, ns1);
    if (rc) return RPBA_E_NOMEM;
    k_set_last<<<1, 1, 0, st>>>((int *)slab_cnt, nslab, 0);
    {
        size_t bytes = 0;
        DCHK(rocprim::exclusive_scan(nullptr, bytes, (const int *)slab_cnt, slab_bs, 0, (size_t)nslab + 1,
                                     rocprim::plus<int>(), st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::exclusive_scan(tmp, bytes, (const int *)slab_cnt, slab_bs, 0, (size_t)nslab + 1,
                                     rocprim::plus<int>(), st));
    }
    if (nslab > 0) {
        k_slab_info<<<nb_(nslab, 256), 256, 0, st>>>(nslab, (unsigned)nc, slab_key, slab_bs, sperm, bk, slab_l,
                                                     slab_mask, slot_slab);
        k_invert<<<nb_(nblk, 256), 256, 0, st>>>(nblk, sperm, sslot);
    }
    k_lower_bound<<<nb_(nstrip + 1, 256), 256, 0, st>>>(nstrip, (unsigned)nc, slab_key, nslab, strip_s0);
    // chunks: about three per SpMV wavefront, 4 to 64 slabs each
    slab_chunk = 8;
    if (spmv_waves > 0) slab_chunk = std::min(64, std::max(4, nslab / (3 * spmv_waves)));
    const int T = slab_chunk;
    k_chunk_counts<<<nb_(nstrip + 1, 256), 256, 0, st>>>(nstrip, T, strip_s0, i32a);
    {
        size_t bytes = 0;
        DCHK(rocprim::exclusive_scan(nullptr, bytes, i32a, chunk_off, 0, (size_t)nstrip + 1, rocprim::plus<int>(), st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::exclusive_scan(tmp, bytes, i32a, chunk_off, 0, (size_t)nstrip + 1, rocprim::plus<int>(), st));
    }
    DCHK(hipMemcpyAsync(hpin, chunk_off + nstrip, sizeof(int), hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    nchunk = hpin[0];
    if (grow(chunk_strip, chunk_strip_cap, (size_t)std::max(nchunk, 1))) return RPBA_E_NOMEM;
    k_chunk_fill<<<nb_(nstrip, 256), 256, 0, st>>>(nstrip, chunk_off, chunk_strip);
    // slabs of each column, strip order (stable sort of slab ids by column)
    if (nslab > 0) {
        if (grow(key_in, key_in_cap, ns1) || grow(key_out, key_out_cap, ns1) || grow(i32b, i32b_cap, ns1))
            return RPBA_E_NOMEM;
        k_iota<<<nb_(nslab, 256), 256, 0, st>>>(nslab, i32b);
        DCHK(hipMemcpyAsync(key_in, slab_l, sizeof(int) * (size_t)nslab, hipMemcpyDeviceToDevice, st));
        const unsigned cb = bits_for((unsigned long long)nc);
        size_t bytes = 0;
        DCHK(rocprim::radix_sort_pairs(nullptr, bytes, key_in, key_out, i32b, cs_list, nslab, 0u, cb, st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::radix_sort_pairs(tmp, bytes, key_in, key_out, i32b, cs_list, nslab, 0u, cb, st));
    }
    k_lower_bound<<<nb_(nc + 1, 256), 256, 0, st>>>(nc, 1u, key_out, nslab, cs_off);
    return RPBA_OK;
}

// ---------------------------------------------------------------------------
// RobustDelete (rbacore.cpp:922-995)
// ---------------------------------------------------------------------------

__global__ void k_rd_mark(int npts, int min_keep, const int *toff, const unsigned char *flag, int *keep_pt, int *newt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > npts) return;
    if (i == npts) {
        keep_pt[i] = 0;
        newt[i] = 0;
        return;
    }
    const int b = toff[i], e = toff[i + 1];
    int nf = 0;
    for (int q = b; q < e; ++q) nf += flag[q] ? 1 : 0;
    const bool drop = nf > 0 && (e - b) - nf < min_keep;
    keep_pt[i] = drop ? 0 : 1;
    newt[i] = drop ? 0 : (e - b) - nf;
}

struct RdArgs {
    int npts;
    const int *toff, *tcam, *tfo, *nrpv, *newi, *newoff;
    const unsigned char *flag;
    const float *ox, *oy, *wx, *wy, *wxy;
    const double *X;
    int *toff2, *tcam2, *tpt2, *tfo2, *nrpv2;
    float *ox2, *oy2, *wx2, *wy2, *wxy2;
    double *X2;
    int *emap;  // old entry -> new entry, or -1
};

// pair p survives when both entries do; its entries renumbered.  flag[P] = 0
// closes the scan.
__global__ void k_remap_pairs(int P, const unsigned long long *pab, const int *emap, unsigned long long *val, int *flag)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > P) return;
    if (p == P) {
        flag[p] = 0;
        return;
    }
    const unsigned long long ab = pab[p];
    const int na = emap[ab >> 32], nb = emap[ab & 0xffffffffull];
    const bool keep = na >= 0 && nb >= 0;
    flag[p] = keep ? 1 : 0;
    val[p] = keep ? (((unsigned long long)na << 32) | (unsigned)nb) : 0ull;
}

__global__ void k_compact_pairs(int P, const unsigned long long *val, const int *flag, const int *pos, unsigned long long *pab)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P && flag[p]) pab[pos[p]] = val[p];
}

// off[b] = pos[off[b]] for b = 0..n (offsets into a list that was compacted)
__global__ void k_remap_offsets(int n, int *off, const int *pos)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= n) off[b] = pos[off[b]];
}

__global__ void k_remap_camq(int n, const int *camq, const int *emap, int *flag)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    flag[j] = (j < n && emap[camq[j]] >= 0) ? 1 : 0;
}

__global__ void k_compact_camq(int n, const int *camq, const int *emap, const int *flag, const int *pos, int *out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n && flag[j]) out[pos[j]] = emap[camq[j]];
}

__global__ void k_rd_scatter(RdArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > a.npts) return;
    if (i == a.npts) {
        a.toff2[a.newi[i]] = a.newoff[i];
        return;
    }
    if (a.newi[i + 1] == a.newi[i]) {  // point erased
        for (int q = a.toff[i]; q < a.toff[i + 1]; ++q) a.emap[q] = -1;
        return;
    }
    const int j = a.newi[i];
    a.toff2[j] = a.newoff[i];
    a.nrpv2[j] = a.nrpv[i];
    for (int c = 0; c < 3; ++c) a.X2[3 * (size_t)j + c] = a.X[3 * (size_t)i + c];
    int w = a.newoff[i];
    for (int q = a.toff[i]; q < a.toff[i + 1]; ++q) {
        if (a.flag[q]) {
            a.emap[q] = -1;
            continue;
        }
        a.emap[q] = w;
        a.tcam2[w] = a.tcam[q];
        a.tpt2[w] = j;
        a.tfo2[w] = a.tfo[q];
        a.ox2[w] = a.ox[q];
        a.oy2[w] = a.oy[q];
        a.wx2[w] = a.wx[q];
        a.wy2[w] = a.wy[q];
        a.wxy2[w] = a.wxy[q];
        ++w;
    }
}

int DevStruct::robust_delete(const unsigned char *flag, int min_keep, double *X, hipStream_t st)
{
    int rc = 0;
    rc |= grow(i32a, i32a_cap, (size_t)std::max(ntr, npts) + 1);
    rc |= grow(i32b, i32b_cap, (size_t)npts + 1);
    rc |= grow(newi, newi_cap, (size_t)npts + 1);
    rc |= grow(newoff, newoff_cap, (size_t)npts + 1);
    rc |= grow(emap, emap_cap, (size_t)std::max(ntr, 1));
    if (rc) return RPBA_E_NOMEM;
    k_rd_mark<<<nb_(npts + 1, 256), 256, 0, st>>>(npts, min_keep, toff, flag, i32a, i32b);
    size_t bytes = 0;
    DCHK(rocprim::exclusive_scan(nullptr, bytes, i32a, newi, 0, (size_t)npts + 1, rocprim::plus<int>(), st));
    if (scratch(bytes)) return RPBA_E_NOMEM;
    DCHK(rocprim::exclusive_scan(tmp, bytes, i32a, newi, 0, (size_t)npts + 1, rocprim::plus<int>(), st));
    DCHK(rocprim::exclusive_scan(nullptr, bytes, i32b, newoff, 0, (size_t)npts + 1, rocprim::plus<int>(), st));
    if (scratch(bytes)) return RPBA_E_NOMEM;
    DCHK(rocprim::exclusive_scan(tmp, bytes, i32b, newoff, 0, (size_t)npts + 1, rocprim::plus<int>(), st));
    if (pinned(2)) return RPBA_E_NOMEM;
    DCHK(hipMemcpyAsync(hpin, newi + npts, sizeof(int), hipMemcpyDeviceToHost, st));
    DCHK(hipMemcpyAsync(hpin + 1, newoff + npts, sizeof(int), hipMemcpyDeviceToHost, st));
    // the flagged entries of kept points are dropped; for a point without any
    // flag every entry is kept (the flag array is zero there).  The output goes
    // to the spare buffers (full initial capacity), which then become the live
    // tracks; the old live buffers are the next spare.
    Tracks &b = spare;
    if (!b.toff && b.alloc(sizes_cap_pts, sizes_cap_tr)) return RPBA_E_NOMEM;
    RdArgs a;
    a.npts = npts; a.toff = toff; a.tcam = tcam; a.tfo = tfo; a.nrpv = nrpv; a.newi = newi; a.newoff = newoff;
    a.flag = flag; a.ox = ox; a.oy = oy; a.wx = w0x; a.wy = w0y; a.wxy = w0xy; a.X = X;
    a.toff2 = b.toff; a.tcam2 = b.tcam; a.tpt2 = b.tpt; a.tfo2 = b.tfo; a.nrpv2 = b.nrpv;
    a.ox2 = b.ox; a.oy2 = b.oy; a.wx2 = b.w0x; a.wy2 = b.w0y; a.wxy2 = b.w0xy; a.X2 = b.X;
    a.emap = emap;
    k_rd_scatter<<<nb_(npts + 1, 256), 256, 0, st>>>(a);
    DCHK(hipStreamSynchronize(st));
    const int np2 = hpin[0], nt2 = hpin[1];
    DCHK(hipMemcpyAsync(X, b.X, sizeof(double) * 3 * (size_t)np2, hipMemcpyDeviceToDevice, st));
    swap_tracks(b);
    const int ntr_old = ntr;
    npts = np2;
    ntr = nt2;
    return update_after_delete(ntr_old, st);
}

// The structure follows the deletion without a rebuild.  Entries keep their
// relative order (points in order, a point's entries in order), so the pair
// list is still grouped by block and in point order inside a block once the
// pairs that lost an entry are dropped, and the camera-major list is still
// sorted once its dropped entries are removed; a rebuild would produce the
// same lists.  A block left without pairs stays as a zero block.
int DevStruct::update_after_delete(int ntr_old, hipStream_t st)
{
    int rc = 0;
    const int P = (int)npairs;
    const size_t n = (size_t)std::max(P, ntr_old) + 1;
    rc |= grow(sflag, sflag_cap, n);
    rc |= grow(spos, spos_cap, n);
    rc |= grow(cq2, cq2_cap, (size_t)std::max(ntr, 1));
    if (rc) return RPBA_E_NOMEM;
    size_t bytes = 0;
    if (pinned(2)) return RPBA_E_NOMEM;
    hpin[0] = 0;
    if (with_pairs) {
        // pairs
        k_remap_pairs<<<nb_(P + 1, 256), 256, 0, st>>>(P, pab, emap, val_in, sflag);
        DCHK(rocprim::exclusive_scan(nullptr, bytes, sflag, spos, 0, (size_t)P + 1, rocprim::plus<int>(), st));
        if (scratch(bytes)) return RPBA_E_NOMEM;
        DCHK(rocprim::exclusive_scan(tmp, bytes, sflag, spos, 0, (size_t)P + 1, rocprim::plus<int>(), st));
        if (P > 0) k_compact_pairs<<<nb_(P, 256), 256, 0, st>>>(P, val_in, sflag, spos, pab);
        k_remap_offsets<<<nb_(nblk + 1, 256), 256, 0, st>>>(nblk, poff, spos);
        DCHK(hipMemcpyAsync(hpin, spos + P, sizeof(int), hipMemcpyDeviceToHost, st));
    }
    // camera-major order
    k_remap_camq<<<nb_(ntr_old + 1, 256), 256, 0, st>>>(ntr_old, camq, emap, sflag);
    DCHK(rocprim::exclusive_scan(nullptr, bytes, sflag, spos, 0, (size_t)ntr_old + 1, rocprim::plus<int>(), st));
    if (scratch(bytes)) return RPBA_E_NOMEM;
    DCHK(rocprim::exclusive_scan(tmp, bytes, sflag, spos, 0, (size_t)ntr_old + 1, rocprim::plus<int>(), st));
    if (ntr_old > 0) k_compact_camq<<<nb_(ntr_old, 256), 256, 0, st>>>(ntr_old, camq, emap, sflag, spos, cq2);
    k_remap_offsets<<<nb_(nc + 1, 256), 256, 0, st>>>(nc, camoff, spos);
    DCHK(hipStreamSynchronize(st));
    npairs = hpin[0];
    std::swap(camq, cq2);
    std::swap(camq_cap, cq2_cap);
    if (ntr > 0) k_invert<<<nb_(ntr, 256), 256, 0, st>>>(ntr, camq, camrank);
    if (point_groups(st)) return RPBA_E_NOMEM;
    rc = chunks_from_camoff(st, &updated_empty);
    if (rc) return rc;
    updated = true;
    return RPBA_OK;
}

int Tracks::alloc(size_t np, size_t nt)
{
    size_t c = 0;
    int rc = 0;
    rc |= grow(toff, c, np + 1);
    c = 0; rc |= grow(nrpv, c, np);
    c = 0; rc |= grow(X, c, 3 * np);
    c = 0; rc |= grow(tcam, c, nt);
    c = 0; rc |= grow(tpt, c, nt);
    c = 0; rc |= grow(tfo, c, nt);
    c = 0; rc |= grow(ox, c, nt);
    c = 0; rc |= grow(oy, c, nt);
    c = 0; rc |= grow(w0x, c, nt);
    c = 0; rc |= grow(w0y, c, nt);
    c = 0; rc |= grow(w0xy, c, nt);
    return rc;
}

void Tracks::release()
{
    hipFree(toff); hipFree(nrpv); hipFree(X); hipFree(tcam); hipFree(tpt); hipFree(tfo);
    hipFree(ox); hipFree(oy); hipFree(w0x); hipFree(w0y); hipFree(w0xy);
    toff = nrpv = tcam = tpt = tfo = nullptr;
    ox = oy = w0x = w0y = w0xy = nullptr;
    X = nullptr;
}

void DevStruct::swap_tracks(Tracks &b)
{
    std::swap(toff, b.toff); std::swap(nrpv, b.nrpv); std::swap(tcam, b.tcam); std::swap(tpt, b.tpt);
    std::swap(tfo, b.tfo); std::swap(ox, b.ox); std::swap(oy, b.oy); std::swap(w0x, b.w0x);
    std::swap(w0y, b.w0y); std::swap(w0xy, b.w0xy);
}

// ---------------------------------------------------------------------------
// PointXXvIn on the device (rbacore.cpp:1918-1974)
// ---------------------------------------------------------------------------

// per observation of the block: the point as a sort key (n_pts for an absent
// observation, image < 0, so it sorts last), its file index as the value, the
// point's count, and the index checks
__global__ __launch_bounds__(256) void k_px_keys(int no, int nc, int np, const int *cam, const int *pt,
                                                 unsigned *key, int *val, int *cnt, int *bad)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < no; i += gridDim.x * blockDim.x) {
        const int c = cam[i], p = pt[i];
        unsigned k = (unsigned)np;
        if (c >= 0) {
            if (c >= nc || p < 0 || p >= np)
                atomicOr(bad, 1);
            else {
                k = (unsigned)p;
                atomicAdd(cnt + p, 1);
            }
        }
        key[i] = k;
        val[i] = i;
    }
}

// per point: a camera twice in its track is an error (checked for every
// point, as the host reading does); kept = |Z| > 1e-8 and at least two
// observations (rbacore.cpp:1954-1955); kept entry count for the scan
__global__ __launch_bounds__(256) void k_px_points(int np, const int *off, const int *sval, const int *cam,
                                                   const float *pts, int *keepp, int *keept, int *bad)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        const int b = off[i], e = off[i + 1], t = e - b;
        for (int x = b; x < e; ++x) {
            const int cx = cam[sval[x]];
            for (int y = x + 1; y < e; ++y)
                if (cam[sval[y]] == cx) atomicOr(bad, 2);
        }
        const bool keep = fabs((double)pts[3 * (size_t)i + 2]) > 1.e-8 && t >= 2;
        keepp[i] = keep ? 1 : 0;
        keept[i] = keep ? t : 0;
    }
}

// kept points: original index, offsets, coordinates; their entries: image,
// kept point, file index, observation and original weights
__global__ __launch_bounds__(256) void k_px_scatter(int np, const int *off, const int *sval, const int *keepp,
                                                    const int *pnew, const int *tnew, const int *cam, const float *xy,
                                                    const float *w, const float *pts, int *toff, int *nrpv, int *tcam,
                                                    int *tpt, int *tfo, float *ox, float *oy, float *w0x, float *w0y,
                                                    float *w0xy, double *X0, int npk_total, int ntr_total)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        if (!keepp[i]) continue;
        const int k = pnew[i], q0 = tnew[i];
        nrpv[k] = i;
        toff[k] = q0;
        if (k == npk_total - 1) toff[k + 1] = ntr_total;
        for (int j = 0; j < 3; ++j) X0[3 * (size_t)k + j] = (double)pts[3 * (size_t)i + j];
        const int b = off[i], e = off[i + 1];
        for (int x = b; x < e; ++x) {
            const int fo = sval[x], q = q0 + (x - b);
            tcam[q] = cam[fo];
            tpt[q] = k;
            tfo[q] = fo;
            ox[q] = xy[2 * (size_t)fo];
            oy[q] = xy[2 * (size_t)fo + 1];
            w0x[q] = w ? w[3 * (size_t)fo] : 1.f;
            w0y[q] = w ? w[3 * (size_t)fo + 1] : 1.f;
            w0xy[q] = w ? w[3 * (size_t)fo + 2] : 0.f;
        }
    }
}

static inline unsigned px_grid(long long n) { return (unsigned)std::min<long long>((n + 255) / 256 + 1, 8192); }

int DevStruct::build_tracks_device(const rpba_block *io, double *X0, hipStream_t st)
{
    const int nc = io->n_cams, np = io->n_pts, no = io->n_obs;
    this->nc = nc;
    // the block's file-order arrays, uploaded as they are
    int *d_cam = nullptr, *d_pt = nullptr, *sval = nullptr, *val = nullptr, *cnt = nullptr, *off = nullptr;
    int *keepp = nullptr, *keept = nullptr, *pnew = nullptr, *tnew = nullptr, *bad = nullptr;
    unsigned *key = nullptr, *skey = nullptr;
    float *d_xy = nullptr, *d_w = nullptr, *d_pts = nullptr;
    void *tmpb = nullptr;
    int rc = RPBA_OK;
    auto fail = [&](int r) { rc = r; };
    const size_t n1 = (size_t)std::max(no, 1), p1 = (size_t)np + 1;
    if (hipMalloc((void **)&d_cam, sizeof(int) * n1) != hipSuccess || hipMalloc((void **)&d_pt, sizeof(int) * n1) != hipSuccess ||
        hipMalloc((void **)&d_xy, sizeof(float) * 2 * n1) != hipSuccess ||
        (io->obs_w && hipMalloc((void **)&d_w, sizeof(float) * 3 * n1) != hipSuccess) ||
        hipMalloc((void **)&d_pts, sizeof(float) * 3 * p1) != hipSuccess || hipMalloc((void **)&key, sizeof(unsigned) * n1) != hipSuccess ||
        hipMalloc((void **)&skey, sizeof(unsigned) * n1) != hipSuccess || hipMalloc((void **)&val, sizeof(int) * n1) != hipSuccess ||
        hipMalloc((void **)&sval, sizeof(int) * n1) != hipSuccess || hipMalloc((void **)&cnt, sizeof(int) * (p1 + 1)) != hipSuccess ||
        hipMalloc((void **)&off, sizeof(int) * (p1 + 1)) != hipSuccess || hipMalloc((void **)&keepp, sizeof(int) * (p1 + 1)) != hipSuccess ||
        hipMalloc((void **)&keept, sizeof(int) * (p1 + 1)) != hipSuccess || hipMalloc((void **)&pnew, sizeof(int) * (p1 + 1)) != hipSuccess ||
        hipMalloc((void **)&tnew, sizeof(int) * (p1 + 1)) != hipSuccess || hipMalloc((void **)&bad, sizeof(int)) != hipSuccess)
        fail(RPBA_E_NOMEM);
    if (!rc && pinned(4)) fail(RPBA_E_NOMEM);
    if (!rc) {
        if (no) {
            if (hipMemcpyAsync(d_cam, io->obs_cam, sizeof(int) * (size_t)no, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(d_pt, io->obs_pt, sizeof(int) * (size_t)no, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(d_xy, io->obs_xy, sizeof(float) * 2 * (size_t)no, hipMemcpyHostToDevice, st) != hipSuccess ||
                (io->obs_w && hipMemcpyAsync(d_w, io->obs_w, sizeof(float) * 3 * (size_t)no, hipMemcpyHostToDevice, st) != hipSuccess))
                fail(RPBA_E_HIP);
        }
        if (!rc && np && hipMemcpyAsync(d_pts, io->pts, sizeof(float) * 3 * (size_t)np, hipMemcpyHostToDevice, st) != hipSuccess)
            fail(RPBA_E_HIP);
    }
    if (!rc && (hipMemsetAsync(cnt, 0, sizeof(int) * (p1 + 1), st) != hipSuccess || hipMemsetAsync(bad, 0, sizeof(int), st) != hipSuccess))
        fail(RPBA_E_HIP);
    if (!rc && no) k_px_keys<<<px_grid(no), 256, 0, st>>>(no, nc, np, d_cam, d_pt, key, val, cnt, bad);
    // stable sort of the file indices by point (file order within a point)
    size_t bytes = 0;
    const unsigned kb = bits_for((unsigned long long)np + 1);
    if (!rc && no) {
        if (rocprim::radix_sort_pairs(nullptr, bytes, key, skey, val, sval, (size_t)no, 0u, kb, st) != hipSuccess) fail(RPBA_E_HIP);
        if (!rc && hipMalloc(&tmpb, std::max<size_t>(bytes, 1)) != hipSuccess) fail(RPBA_E_NOMEM);
        if (!rc && rocprim::radix_sort_pairs(tmpb, bytes, key, skey, val, sval, (size_t)no, 0u, kb, st) != hipSuccess)
            fail(RPBA_E_HIP);
        if (tmpb) (void)hipFree(tmpb);
        tmpb = nullptr;
    }
    // per-point offsets into the sorted list, then the kept points
    auto scan = [&](const int *in, int *out, size_t n) -> int {
        size_t b2 = 0;
        if (rocprim::exclusive_scan(nullptr, b2, in, out, 0, n, rocprim::plus<int>(), st) != hipSuccess) return RPBA_E_HIP;
        void *t2 = nullptr;
        if (hipMalloc(&t2, std::max<size_t>(b2, 1)) != hipSuccess) return RPBA_E_NOMEM;
        const int r2 = rocprim::exclusive_scan(t2, b2, in, out, 0, n, rocprim::plus<int>(), st) != hipSuccess ? (int)RPBA_E_HIP : (int)RPBA_OK;
        (void)hipFree(t2);
        return r2;
    };
    if (!rc) rc = scan(cnt, off, p1);
    if (!rc && np) k_px_points<<<px_grid(np), 256, 0, st>>>(np, off, sval, d_cam, d_pts, keepp, keept, bad);
    if (!rc && hipMemsetAsync(keepp + np, 0, sizeof(int), st) != hipSuccess) fail(RPBA_E_HIP);
    if (!rc && hipMemsetAsync(keept + np, 0, sizeof(int), st) != hipSuccess) fail(RPBA_E_HIP);
    if (!rc) rc = scan(keepp, pnew, p1);
    if (!rc) rc = scan(keept, tnew, p1);
    if (!rc) {
        if (hipMemcpyAsync(hpin, pnew + np, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hpin + 1, tnew + np, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hpin + 2, bad, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            fail(RPBA_E_HIP);
    }
    if (!rc && hpin[2]) fail(RPBA_E_INVALID);
    if (!rc) {
        npts = hpin[0];
        ntr = hpin[1];
        sizes_cap_pts = (size_t)npts;
        sizes_cap_tr = (size_t)ntr;
        Tracks t;
        if (!toff) {
            if (t.alloc(sizes_cap_pts, sizes_cap_tr)) fail(RPBA_E_NOMEM);
            if (!rc) {
                toff = t.toff; nrpv = t.nrpv; tcam = t.tcam; tpt = t.tpt; tfo = t.tfo;
                ox = t.ox; oy = t.oy; w0x = t.w0x; w0y = t.w0y; w0xy = t.w0xy;
                (void)hipFree(t.X);
            }
        }
    }
    if (!rc && hipMemsetAsync(toff, 0, sizeof(int), st) != hipSuccess) fail(RPBA_E_HIP);
    if (!rc && np)
        k_px_scatter<<<px_grid(np), 256, 0, st>>>(np, off, sval, keepp, pnew, tnew, d_cam, d_xy, d_w, d_pts, toff, nrpv, tcam,
                                                  tpt, tfo, ox, oy, w0x, w0y, w0xy, X0, npts, ntr);
    if (!rc && hipGetLastError() != hipSuccess) fail(RPBA_E_HIP);
    if (!rc) {
        // snapshot for restore_tracks
        npts0 = npts;
        ntr0 = ntr;
        if (!init.toff && init.alloc(sizes_cap_pts, sizes_cap_tr)) fail(RPBA_E_NOMEM);
    }
    if (!rc) {
        const size_t pn = (size_t)npts, nt = (size_t)ntr;
        if (hipMemcpyAsync(init.toff, toff, sizeof(int) * (pn + 1), hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.nrpv, nrpv, sizeof(int) * pn, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.tcam, tcam, sizeof(int) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.tpt, tpt, sizeof(int) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.tfo, tfo, sizeof(int) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.ox, ox, sizeof(float) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.oy, oy, sizeof(float) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.w0x, w0x, sizeof(float) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.w0y, w0y, sizeof(float) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(init.w0xy, w0xy, sizeof(float) * nt, hipMemcpyDeviceToDevice, st) != hipSuccess)
            fail(RPBA_E_HIP);
    }
    if (hipStreamSynchronize(st) != hipSuccess && !rc) fail(RPBA_E_HIP);
    for (void *pp : {(void *)d_cam, (void *)d_pt, (void *)d_xy, (void *)d_w, (void *)d_pts, (void *)key, (void *)skey,
                     (void *)val, (void *)sval, (void *)cnt, (void *)off, (void *)keepp, (void *)keept, (void *)pnew,
                     (void *)tnew, (void *)bad})
        if (pp) (void)hipFree(pp);
    return rc;
}

int DevStruct::upload_tracks(const Structure &s, const float *obs_xy, const std::vector<float> &w0, hipStream_t st)
{
    nc = s.nc;
    npts = s.npts;
    ntr = s.ntr;
    sizes_cap_pts = (size_t)npts;
    sizes_cap_tr = (size_t)ntr;
    Tracks t;
    if (!toff || (size_t)npts > trk_np_alloc || (size_t)ntr > trk_nt_alloc) {
        // first upload, or a structure larger than the room (a solve that
        // uploads again, as the partitioned one does per outer iteration,
        // keeps its buffers otherwise)
        if (toff) {
            Tracks old;
            old.toff = toff; old.nrpv = nrpv; old.tcam = tcam; old.tpt = tpt; old.tfo = tfo;
            old.ox = ox; old.oy = oy; old.w0x = w0x; old.w0y = w0y; old.w0xy = w0xy;
            old.release();
            init.release();
        }
        if (t.alloc(sizes_cap_pts, sizes_cap_tr)) return RPBA_E_NOMEM;
        toff = t.toff; nrpv = t.nrpv; tcam = t.tcam; tpt = t.tpt; tfo = t.tfo;
        ox = t.ox; oy = t.oy; w0x = t.w0x; w0y = t.w0y; w0xy = t.w0xy;
        (void)hipFree(t.X);
        trk_np_alloc = sizes_cap_pts;
        trk_nt_alloc = sizes_cap_tr;
    }
    // every array through one pinned staging buffer (grown when needed), so
    // the copies run at the link's rate; filled over ranges of entries
    const size_t np1 = (size_t)npts + 1, nt = (size_t)ntr;
    const size_t need = sizeof(int) * (2 * np1 + 8 * nt);
    if (need > stage_cap) {
        if (stage) (void)hipHostFree(stage);
        stage = nullptr;
        stage_cap = 0;
        if (hipHostMalloc(&stage, need, hipHostMallocDefault) != hipSuccess) return RPBA_E_NOMEM;
        stage_cap = need;
    }
    int *h_toff = (int *)stage, *h_nrpv = h_toff + np1, *h_tcam = h_nrpv + np1, *h_tpt = h_tcam + nt;
    int *h_tfo = h_tpt + nt;
    float *hx = (float *)(h_tfo + nt), *hy = hx + nt, *a = hy + nt, *b = a + nt, *c = b + nt;
    std::memcpy(h_toff, s.toff.data(), sizeof(int) * np1);
    std::memcpy(h_nrpv, s.nrpv.data(), sizeof(int) * (size_t)npts);
    parallel_ranges(ntr, host_threads(), [&](long long q0, long long q1, int) {
        for (long long q = q0; q < q1; ++q) {
            const int fo = s.tfo[q];
            h_tcam[q] = s.tcam[q];
            h_tpt[q] = s.tpt[q];
            h_tfo[q] = fo;
            hx[q] = obs_xy[2 * (size_t)fo];
            hy[q] = obs_xy[2 * (size_t)fo + 1];
            a[q] = w0[3 * (size_t)q];
            b[q] = w0[3 * (size_t)q + 1];
            c[q] = w0[3 * (size_t)q + 2];
        }
    });
    DCHK(hipMemcpyAsync(toff, h_toff, sizeof(int) * np1, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(nrpv, h_nrpv, sizeof(int) * (size_t)npts, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(tcam, h_tcam, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(tpt, h_tpt, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(tfo, h_tfo, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(ox, hx, sizeof(float) * nt, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(oy, hy, sizeof(float) * nt, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(w0x, a, sizeof(float) * nt, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(w0y, b, sizeof(float) * nt, hipMemcpyHostToDevice, st));
    DCHK(hipMemcpyAsync(w0xy, c, sizeof(float) * nt, hipMemcpyHostToDevice, st));
    // snapshot for restore_tracks
    npts0 = npts;
    ntr0 = ntr;
    if (!init.toff && init.alloc(sizes_cap_pts, sizes_cap_tr)) return RPBA_E_NOMEM;
    const size_t np = (size_t)npts;
    DCHK(hipMemcpyAsync(init.toff, toff, sizeof(int) * (np + 1), hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.nrpv, nrpv, sizeof(int) * np, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.tcam, tcam, sizeof(int) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.tpt, tpt, sizeof(int) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.tfo, tfo, sizeof(int) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.ox, ox, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.oy, oy, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.w0x, w0x, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.w0y, w0y, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(init.w0xy, w0xy, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipStreamSynchronize(st));
    return RPBA_OK;
}

int DevStruct::restore_tracks(hipStream_t st)
{
    updated = false;  // the initial tracks need the full build
    npts = npts0;
    ntr = ntr0;
    const size_t np = (size_t)npts, nt = (size_t)ntr;
    DCHK(hipMemcpyAsync(toff, init.toff, sizeof(int) * (np + 1), hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(nrpv, init.nrpv, sizeof(int) * np, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(tcam, init.tcam, sizeof(int) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(tpt, init.tpt, sizeof(int) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(tfo, init.tfo, sizeof(int) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(ox, init.ox, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(oy, init.oy, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(w0x, init.w0x, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(w0y, init.w0y, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    DCHK(hipMemcpyAsync(w0xy, init.w0xy, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    return RPBA_OK;
}

int DevStruct::download_tracks(std::vector<int> &toff_h, std::vector<int> &nrpv_h, std::vector<int> &tfo_h,
                               std::vector<int> &tcam_h, hipStream_t st) const
{
    toff_h.resize((size_t)npts + 1);
    nrpv_h.resize((size_t)npts);
    tfo_h.resize((size_t)ntr);
    tcam_h.resize((size_t)ntr);
    DCHK(hipMemcpyAsync(toff_h.data(), toff, sizeof(int) * toff_h.size(), hipMemcpyDeviceToHost, st));
    if (npts) DCHK(hipMemcpyAsync(nrpv_h.data(), nrpv, sizeof(int) * nrpv_h.size(), hipMemcpyDeviceToHost, st));
    if (ntr) DCHK(hipMemcpyAsync(tfo_h.data(), tfo, sizeof(int) * tfo_h.size(), hipMemcpyDeviceToHost, st));
    if (ntr) DCHK(hipMemcpyAsync(tcam_h.data(), tcam, sizeof(int) * tcam_h.size(), hipMemcpyDeviceToHost, st));
    DCHK(hipStreamSynchronize(st));
    return RPBA_OK;
}

void DevStruct::release()
{
    init.release();
    spare.release();
    if (stage) (void)hipHostFree(stage);
    stage = nullptr;
    stage_cap = 0;
    trk_np_alloc = trk_nt_alloc = 0;
    if (hpin) (void)hipHostFree(hpin);
    hpin = nullptr;
    hpin_cap = 0;
    Tracks t;
    t.toff = toff; t.nrpv = nrpv; t.tcam = tcam; t.tpt = tpt; t.tfo = tfo;
    t.ox = ox; t.oy = oy; t.w0x = w0x; t.w0y = w0y; t.w0xy = w0xy; t.X = nullptr;
    t.release();
    toff = nrpv = tcam = tpt = tfo = nullptr;
    ox = oy = w0x = w0y = w0xy = nullptr;
    auto fr = [](auto *&p, size_t &c) {
        if (p) (void)hipFree(p);
        p = nullptr;
        c = 0;
    };
    fr(camoff, camoff_cap); fr(camq, camq_cap); fr(camrank, camrank_cap); fr(pgrp, pgrp_cap); fr(pg4, pg4_cap); fr(ch4, ch4_cap); fr(ch_start, ch_start_cap); fr(ch_off, ch_off_cap); fr(u32a, u32a_cap); fr(i32a, i32a_cap); fr(i32b, i32b_cap);
    fr(ppoff, ppoff_cap); fr(key_in, key_in_cap); fr(key_out, key_out_cap); fr(val_in, val_in_cap);
    fr(sslot, sslot_cap); fr(sperm, sperm_cap); fr(slot_slab, slot_slab_cap); fr(slab_bs, slab_bs_cap);
    fr(slab_l, slab_l_cap); fr(slab_mask, slab_mask_cap); fr(strip_s0, strip_s0_cap); fr(chunk_off, chunk_off_cap);
    fr(chunk_strip, chunk_strip_cap); fr(cs_off, cs_off_cap); fr(cs_list, cs_list_cap);
    fr(slab_key, slab_key_cap); fr(slab_cnt, slab_cnt_cap);
    fr(pab, pab_cap); fr(ukey, ukey_cap); fr(ucnt, ucnt_cap); fr(nruns, nruns_cap); fr(poff, poff_cap);
    fr(bk, bk_cap); fr(bl, bl_cap); fr(roff, roff_cap); fr(coff, coff_cap); fr(cblk, cblk_cap);
    fr(key2, key2_cap); fr(key2s, key2s_cap); fr(bid, bid_cap); fr(newi, newi_cap); fr(newoff, newoff_cap);
    fr(tmp, tmp_cap); fr(emap, emap_cap); fr(sflag, sflag_cap); fr(spos, spos_cap); fr(cq2, cq2_cap);
}

}  // namespace rpba
