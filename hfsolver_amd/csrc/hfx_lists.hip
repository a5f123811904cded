// Schwarz-screened quartet lists built on the GPU (gfx950).
//
// From the shell pairs and their Schwarz bounds Q (computed on the GPU by
// schwarz_kernel) this part builds, per geometry:
//   - the significant pairs (Q_ab Q_max >= threshold) of every pair class,
//     Q-descending (stable radix sorts: Q, then the class), and the same pairs
//     grouped into column windows and, within those, sublists by primitive-pair
//     count bucket (a third, stable sort on class, window and bucket), with run
//     lengths of the (class, window, bucket) keys;
//   - every class's ket table (ket_fill_kernel: records, C - D planes, chunked
//     primitive planes in FP64 and FP32, and per ket its bound, class position
//     and primitive count), the per-sublist prefix sums of primitive counts;
//   - for every class combination and bra row, each sublist's ket range with
//     Q_ab Q_cd >= threshold, split at the FP32 cut (binary searches in the
//     Q-descending sublists; same-class combinations stop at the diagonal),
//     and the quartet and primitive-quartet counts of the launch;
//   - the row-mode block lists: per group of R bra rows the union of its
//     ranges, cut into chunks that meet at least one range (count, scan, fill).
// The host reads back only counts (class sizes, sublist runs, the summed
// spans of each launch, block totals) to size allocations and choose chunks.
//
// Counterpart of nothing in hfsolver: the reference evaluates every
// basis-function quartet (src/qc.c:95-132) with no screening.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "hfx_internal.h"

namespace {

// Sort key of a bound q >= 0: descending q <-> ascending key.
__device__ HFX_INLINE unsigned long long qdesc_key(double q) {
    return ~(unsigned long long)__double_as_longlong(q);
}

__device__ HFX_INLINE int np_bucket_dev(int np) {
    const int B[14] = {1, 3, 6, 9, 14, 18, 27, 36, 54, 72, 100, 144, 196, 256};
#pragma unroll
    for (int i = 0; i < 14; i++)
        if (np <= B[i]) return B[i];
    return (np + 255) / 256 * 256;
}

// Largest Q of all pairs (as the bit pattern of a non-negative double).
__global__ void qmax_kernel(const ShellPair* __restrict__ pairs, int np, unsigned long long* qmax) {
    unsigned long long m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x)
        m = max(m, (unsigned long long)__double_as_longlong(pairs[i].Q));
    atomicMax(qmax, m);
}

// Sort keys: Q descending for the significant pairs, class (7 = not significant).
__global__ void pair_key_kernel(const ShellPair* __restrict__ pairs, int np, const unsigned long long* qmax,
                                double thr, unsigned long long* kq, unsigned* kc, int* idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double q = pairs[i].Q;
    const double qm = __longlong_as_double((long long)*qmax);
    const bool sig = q * qm >= thr && q > 0;
    kq[i] = sig ? qdesc_key(q) : ~0ull;
    kc[i] = sig ? (unsigned)pairs[i].cls : 7u;
    idx[i] = i;
}

// Class key (order 1) and class-window-bucket key (order 2: class << 24 |
// column window << 16 | bucket) of the pairs in a given order; pwin == nullptr:
// window 0 for every pair.
__global__ void gather_key_kernel(const ShellPair* __restrict__ pairs, int np, const int* __restrict__ order,
                                  const unsigned* __restrict__ kc, const int* __restrict__ pwin, unsigned* kc_out,
                                  unsigned* kb_out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    const int pid = order[j];
    const unsigned c = kc[pid];
    if (kc_out) kc_out[j] = c;
    if (kb_out) {
        const unsigned w = pwin ? (unsigned)pwin[pid] : 0u;
        kb_out[j] = c == 7u ? (7u << 24) : ((c << 24) | (w << 16) | (unsigned)np_bucket_dev(pairs[pid].np));
    }
}

// Position of every significant pair in its class list (order 1, class c at [off[c], off[c+1])).
__global__ void pos_kernel(const int* __restrict__ order, const unsigned* __restrict__ kc_sorted, int nsig,
                           const int* __restrict__ cls_off, int* pos) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsig) return;
    pos[order[j]] = j - cls_off[kc_sorted[j]];
}

// One class's ket table: sublist s = blockIdx.y, ket j = position within it.
__global__ void ket_fill_kernel(KetTab t, KetOut o, KetAux x, const int* __restrict__ sub_order,
                                const ShellPair* __restrict__ pairs, const PrimPair* __restrict__ prims,
                                const PrimExp* __restrict__ pexp, const int* __restrict__ pos) {
    const int s = blockIdx.y;
    const int nks = t.off[s + 1] - t.off[s];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nks) return;
    const int k = t.off[s] + j;
    const int nk = t.nk;
    const int pid = sub_order[k];
    const ShellPair sp = pairs[pid];
    o.info[k] = KetInfo{pid, sp.fa, sp.fb, sp.cpa | (sp.cpb << 15) | ((sp.A == sp.B ? 1 : 0) << 30)};
    o.shcd[k] = (sp.A & 0xffff) | (sp.B << 16);
    o.qk[k] = (float)(sp.Q * (1.0 + 1e-6));
    o.cd[k] = sp.ABx;
    o.cd[(size_t)nk + k] = sp.ABy;
    o.cd[2 * (size_t)nk + k] = sp.ABz;
    x.q[k] = sp.Q;
    x.pos[k] = pos[pid];
    x.np[k] = sp.np;
    // primitive planes (padding primitives: the last one with K = 0), fields in
    // the slot order of KetChunk: p, 1/p, P, K, (P - C)_x, (P - C)_y | (P - C)_z, C exponent
    const int pitch = ket_pitch(nks);
    double* pd = o.prim + t.psoff[s];
    float* pf = o.primf + t.psoff_f[s];
    for (int i = 0; i < t.np[s]; i++) {
        const int pi = sp.p0 + min(i, sp.np - 1);
        const PrimPair q = prims[pi];
        const double v[KETPRIM_FIELDS] = {q.p, q.ip, q.Px, q.Py, q.Pz, i < sp.np ? q.K : 0.0, q.PAx, q.PAy, q.PAz,
                                          pexp[pi].aa};
        const size_t base = (size_t)i * KETPRIM_FIELDS * pitch;
#pragma unroll
        for (int slot = 0; slot < KETPRIM_FIELDS; slot++) {
            const size_t wd = slot < 8 ? base + ((size_t)(slot / 2) * pitch + j) * 2 + slot % 2
                                       : base + (size_t)slot * pitch + j;
            const size_t wf = slot < 8 ? base + ((size_t)(slot / 4) * pitch + j) * 4 + slot % 4
                                       : base + (size_t)slot * pitch + j;
            pd[wd] = v[slot];
            pf[wf] = (float)v[slot];
        }
    }
}

// Number of leading elements of the descending q[0..n) that are >= need.
__device__ HFX_INLINE int count_ge(const double* __restrict__ q, int n, double need) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (q[mid] >= need)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Number of leading elements of the ascending p[0..n) that are <= v.
__device__ HFX_INLINE int count_le(const int* __restrict__ p, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (p[mid] <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Ket ranges of every (bra row i, sublist s) of one class combination: FP64
// share [0, k64), FP32 share [k64, kmax) as ket-table positions; quartet and
// primitive-quartet counts per (sublist, precision) into stats[s][prec][2].
__global__ void row_ranges_kernel(RangeArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    double nq[2] = {0, 0}, npq[2] = {0, 0};
    if (i < a.nb) {
        const int pid = a.bra_list[i];
        const double qi = a.pairs[pid].Q;
        const int npb = a.pairs[pid].np;
        const int soff = a.kt.off[s], nk = a.kt.off[s + 1] - soff;
        const double* kq = a.x.q + soff;
        int km = qi > 0 ? count_ge(kq, nk, a.thr / qi) : 0;
        if (a.same) km = min(km, count_le(a.x.pos + soff, nk, i));
        const int k6 = a.thr32 > 0 ? (qi > 0 ? count_ge(kq, nk, a.thr32 / qi) : 0) : km;
        const int k64 = min(k6, km);
        const int kb[2] = {0, k64}, ke[2] = {k64, km};
        const int* cnp = a.x.cnp + soff;   // prefix sums of primitive counts, from the class start
#pragma unroll
        for (int prec = 0; prec < 2; prec++) {
            const size_t x = (size_t)i * a.S + s;
            a.rkb[prec][x] = soff + kb[prec];
            a.rke[prec][x] = soff + max(kb[prec], ke[prec]);
            if (ke[prec] > kb[prec]) {
                nq[prec] = ke[prec] - kb[prec];
                npq[prec] = (double)npb * (double)(cnp[ke[prec]] - cnp[kb[prec]]);
            }
        }
    }
    // block sums, then one atomic per block and quantity (integer-valued sums: exact)
    __shared__ double red[4][256];
    red[0][threadIdx.x] = nq[0];
    red[1][threadIdx.x] = npq[0];
    red[2][threadIdx.x] = nq[1];
    red[3][threadIdx.x] = npq[1];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++)
            if (red[q][0] != 0.0) atomicAdd(&a.stats[(size_t)s * 4 + q], red[q][0]);
}

// Union [lo, hi) over the rows of group g in sublist s of one precision's ranges.
__device__ HFX_INLINE void group_range(const BlockArgsL& a, int g0, int s, int& lo, int& hi) {
    lo = 1 << 30;
    hi = 0;
    for (int i = g0; i < min(a.nb, g0 + a.R); i++) {
        const size_t x = (size_t)i * a.S + s;
        const int b = a.rkb[x], e = a.rke[x];
        if (e > b) {
            lo = min(lo, b);
            hi = max(hi, e);
        }
    }
}

// Per group of R rows: summed span of its ranges (sums[0]) and non-empty groups (sums[1]).
__global__ void group_span_kernel(BlockArgsL a, double* sums) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    double span = 0;
    if (g * a.R < a.nb)
        for (int s = 0; s < a.S; s++) {
            int lo, hi;
            group_range(a, g * a.R, s, lo, hi);
            if (hi > lo) span += hi - lo;
        }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = span;
    red[1][threadIdx.x] = span > 0 ? 1.0 : 0.0;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (red[0][0] != 0.0) atomicAdd(&sums[0], red[0][0]);
        if (red[1][0] != 0.0) atomicAdd(&sums[1], red[1][0]);
    }
}

// Blocks of group g: chunks [k0, k0 + chunk) from its first range's start to its
// last range's end that meet at least one of its ranges. Counts them (out ==
// nullptr) or writes them from offset off[g].
__global__ void group_blocks_kernel(BlockArgsL a, int* count, const int* off, int* brow, int* bk0) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int g0 = g * a.R;
    if (g0 >= a.nb) return;
    int first = -1, last = 0;
    for (int s = 0; s < a.S; s++) {
        int lo, hi;
        group_range(a, g0, s, lo, hi);
        if (hi > lo) {
            if (first < 0) first = lo;
            last = hi;
        }
    }
    int n = 0;
    if (first >= 0)
        for (int k0 = first; k0 < last; k0 += a.chunk) {
            bool meets = false;
            for (int s = 0; s < a.S && !meets; s++) {
                int lo, hi;
                group_range(a, g0, s, lo, hi);
                meets = hi > lo && lo < k0 + a.chunk && hi > k0;
            }
            if (!meets) continue;
            if (brow) {
                brow[off[g] + n] = g0;
                bk0[off[g] + n] = k0;
            }
            n++;
        }
    if (count) count[g] = n;
}

template <typename T>
hipError_t temp_call(T&& f, hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = f(nullptr, bytes);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    if ((e = hipMallocAsync(&tmp, bytes > 0 ? bytes : 16, s)) != hipSuccess) return e;
    e = f(tmp, bytes);
    const hipError_t e2 = hipFreeAsync(tmp, s);
    return e != hipSuccess ? e : e2;
}

int nblocks(long long n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

hipError_t hfx_pair_order(const ShellPair* d_pairs, const int* d_pwin, int np, double thr, PairOrder& o,
                          hipStream_t s) {
    hipError_t e;
#define HFX_LCHK(x)                    \
    if ((e = (x)) != hipSuccess) {     \
        return e;                      \
    }
    unsigned long long *kq = nullptr, *kq2 = nullptr, *qmax = nullptr;
    unsigned *kc = nullptr, *kc2 = nullptr, *kc3 = nullptr, *kb = nullptr, *kb2 = nullptr, *runkey = nullptr;
    int *idx = nullptr, *idx2 = nullptr, *runcnt = nullptr, *nruns = nullptr, *cls_off = nullptr;
    auto cleanup = [&]() {
        for (void* p : {(void*)kq, (void*)kq2, (void*)qmax, (void*)kc, (void*)kc2, (void*)kc3, (void*)kb, (void*)kb2,
                        (void*)runkey, (void*)idx, (void*)idx2, (void*)runcnt, (void*)nruns, (void*)cls_off})
            if (p) (void)hipFreeAsync(p, s);
    };
    struct Guard {
        decltype(cleanup)& f;
        ~Guard() { f(); }
    } guard{cleanup};
    const size_t n = (size_t)np;
    HFX_LCHK(hipMallocAsync((void**)&kq, n * 8, s));
    HFX_LCHK(hipMallocAsync((void**)&kq2, n * 8, s));
    HFX_LCHK(hipMallocAsync((void**)&qmax, 8, s));
    HFX_LCHK(hipMallocAsync((void**)&kc, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&kc2, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&kc3, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&kb, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&kb2, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&runkey, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&idx, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&idx2, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&runcnt, n * 4, s));
    HFX_LCHK(hipMallocAsync((void**)&nruns, 4, s));
    HFX_LCHK(hipMallocAsync((void**)&cls_off, 8 * 4, s));
    HFX_LCHK(hipMemsetAsync(qmax, 0, 8, s));
    hipLaunchKernelGGL(qmax_kernel, dim3(std::min(nblocks(np), 1024)), dim3(256), 0, s, d_pairs, np, qmax);
    HFX_LCHK(hipGetLastError());
    hipLaunchKernelGGL(pair_key_kernel, dim3(nblocks(np)), dim3(256), 0, s, d_pairs, np, qmax, thr, kq, kc, idx);
    HFX_LCHK(hipGetLastError());
    // order 1: Q descending (stable in the pair index), then the class (stable)
    HFX_LCHK(temp_call([&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortPairs(t, b, kq, kq2, idx, idx2, np, 0, 64, s);
    }, s));
    hipLaunchKernelGGL(gather_key_kernel, dim3(nblocks(np)), dim3(256), 0, s, d_pairs, np, idx2, kc, d_pwin, kc2,
                       nullptr);
    HFX_LCHK(hipGetLastError());
    HFX_LCHK(temp_call([&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortPairs(t, b, kc2, kc3, idx2, o.cls_order, np, 0, 3, s);
    }, s));
    // order 2: class, column window and primitive-count bucket (stable in order 1)
    hipLaunchKernelGGL(gather_key_kernel, dim3(nblocks(np)), dim3(256), 0, s, d_pairs, np, o.cls_order, kc, d_pwin,
                       nullptr, kb);
    HFX_LCHK(hipGetLastError());
    HFX_LCHK(temp_call([&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortPairs(t, b, kb, kb2, o.cls_order, o.sub_order, np, 0, 27, s);
    }, s));
    HFX_LCHK(temp_call([&](void* t, size_t& b) {
        return hipcub::DeviceRunLengthEncode::Encode(t, b, kb2, runkey, runcnt, nruns, np, s);
    }, s));
    int nr = 0;
    HFX_LCHK(hipMemcpyAsync(&nr, nruns, 4, hipMemcpyDeviceToHost, s));
    HFX_LCHK(hipStreamSynchronize(s));
    std::vector<unsigned> hk(nr);
    std::vector<int> hc(nr);
    HFX_LCHK(hipMemcpyAsync(hk.data(), runkey, (size_t)nr * 4, hipMemcpyDeviceToHost, s));
    HFX_LCHK(hipMemcpyAsync(hc.data(), runcnt, (size_t)nr * 4, hipMemcpyDeviceToHost, s));
    HFX_LCHK(hipStreamSynchronize(s));
    o.runs.clear();
    for (int c = 0; c < 8; c++) o.ncls[c] = 0;
    for (int r = 0; r < nr; r++) {
        const int c = (int)(hk[r] >> 24), w = (int)((hk[r] >> 16) & 0xff), b = (int)(hk[r] & 0xffff);
        if (c >= HFX_NPAIRCLASS) continue;
        o.runs.push_back({c, w, b, hc[r]});
        o.ncls[c] += hc[r];
    }
    int off[8] = {0};
    for (int c = 1; c < 8; c++) off[c] = off[c - 1] + (c - 1 < HFX_NPAIRCLASS ? o.ncls[c - 1] : 0);
    o.nsig = off[HFX_NPAIRCLASS];
    HFX_LCHK(hipMemcpyAsync(cls_off, off, sizeof off, hipMemcpyHostToDevice, s));
    HFX_LCHK(hipMemsetAsync(o.pos, 0xff, n * 4, s));
    if (o.nsig > 0) {
        hipLaunchKernelGGL(pos_kernel, dim3(nblocks(o.nsig)), dim3(256), 0, s, o.cls_order, kc3, o.nsig, cls_off, o.pos);
        HFX_LCHK(hipGetLastError());
    }
    HFX_LCHK(hipStreamSynchronize(s));
    return hipSuccess;
#undef HFX_LCHK
}

hipError_t hfx_launch_ket_fill(const KetTab& t, const KetOut& o, const KetAux& x, const int* sub_order,
                               const ShellPair* pairs, const PrimPair* prims, const PrimExp* pexp, const int* pos,
                               hipStream_t s) {
    int maxk = 0;
    for (int i = 0; i < t.S; i++) maxk = std::max(maxk, t.off[i + 1] - t.off[i]);
    if (maxk == 0) return hipSuccess;
    hipLaunchKernelGGL(ket_fill_kernel, dim3(nblocks(maxk), t.S), dim3(256), 0, s, t, o, x, sub_order, pairs, prims,
                       pexp, pos);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // prefix sums of the primitive counts over the class's ket order (x.cnp[0] = 0)
    return temp_call([&](void* tmp, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(tmp, b, x.np, x.cnp, t.nk + 1, s);
    }, s);
}

hipError_t hfx_launch_row_ranges(const RangeArgs& a, hipStream_t s) {
    if (a.nb == 0 || a.S == 0) return hipSuccess;
    hipLaunchKernelGGL(row_ranges_kernel, dim3(nblocks(a.nb), a.S), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t hfx_launch_group_spans(const BlockArgsL& a, double* sums, hipStream_t s) {
    const int ng = (a.nb + a.R - 1) / a.R;
    if (ng == 0) return hipSuccess;
    hipLaunchKernelGGL(group_span_kernel, dim3(nblocks(ng)), dim3(256), 0, s, a, sums);
    return hipGetLastError();
}

hipError_t hfx_group_blocks(const BlockArgsL& a, int** brow, int** bk0, int* nblk, std::vector<void*>& allocs,
                            hipStream_t s) {
    const int ng = (a.nb + a.R - 1) / a.R;
    *nblk = 0;
    if (ng == 0) return hipSuccess;
    int *cnt = nullptr, *off = nullptr;
    hipError_t e = hipMallocAsync((void**)&cnt, (size_t)(ng + 1) * 4, s);
    if (e == hipSuccess) e = hipMallocAsync((void**)&off, (size_t)(ng + 1) * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, (size_t)(ng + 1) * 4, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(group_blocks_kernel, dim3(nblocks(ng)), dim3(256), 0, s, a, cnt, (const int*)nullptr,
                           (int*)nullptr, (int*)nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = temp_call([&](void* tmp, size_t& b) { return hipcub::DeviceScan::ExclusiveSum(tmp, b, cnt, off, ng + 1, s); },
                      s);
    int total = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&total, off + ng, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && total > 0) {
        e = hipMalloc((void**)brow, (size_t)total * 4);
        if (e == hipSuccess) allocs.push_back(*brow);
        if (e == hipSuccess) e = hipMalloc((void**)bk0, (size_t)total * 4);
        if (e == hipSuccess) allocs.push_back(*bk0);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(group_blocks_kernel, dim3(nblocks(ng)), dim3(256), 0, s, a, (int*)nullptr,
                               (const int*)off, *brow, *bk0);
            e = hipGetLastError();
        }
    }
    if (cnt) (void)hipFreeAsync(cnt, s);
    if (off) (void)hipFreeAsync(off, s);
    if (e == hipSuccess) *nblk = total;
    return e;
}
