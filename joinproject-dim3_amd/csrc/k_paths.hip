// Join paths grouped by (x, z): the device side of a caller-written
// aggregate policy.
//
// The reference's join-aggregate takes its aggregate as a policy class
// (SpmmAggregate<..., _Aggregate_type>, dim3.hpp:1263-1311; its shipped
// policy AggregateFunctionAvg, dim3.hpp:1453-1480): per x row it calls
// insert(z, v1, v2) on the first path of each (x, z) group, update(z, v1, v2)
// on the others and get(z) for every group at the row's end, z being a dense
// z ID.  The library's own policies (COUNT, SUM, MIN, MAX, AVG) run inside
// the expansion kernels; a policy the library does not know is the caller's
// own host code, so the device delivers what it needs to run it: every join
// path's R and S value, grouped by (x, z).
//
//   1. S tuple indices sorted by join key (stable radix sort);
//   2. each R tuple's S range by binary search, path offsets by a scan;
//   3. one 64-bit (x, z) key per join path, stable radix sort of the path
//      indices by it (so a group's paths keep (R tuple, S tuple) order: the
//      reference's order within a row, R_all row then S row);
//   4. groups by run-length encoding, a dense z key per group (the rank of z
//      among the distinct S attribute values), the path values gathered in
//      group order.
// Per join path: 8 + 4 bytes of key and index through two sort passes, and
// 2 value gathers.  Paths with x == z are dropped under exclude_equal.
#include <hipcub/hipcub.hpp>

#include <functional>

#include "dim3_internal.h"

namespace {

__device__ __forceinline__ uint32_t flip(int32_t v) { return (uint32_t)v ^ 0x80000000u; }
__device__ __forceinline__ int32_t unflip(uint32_t v) { return (int32_t)(v ^ 0x80000000u); }

inline unsigned blocks_for(uint64_t n) { return (unsigned)(n == 0 ? 1 : (n + 255) / 256); }

// attribute (offset 0) and sortable join key (offset 4) of each raw tuple
__global__ void k_split_raw(const uint8_t* __restrict__ raw, uint64_t n, uint32_t stride, uint32_t* __restrict__ attr,
                            uint32_t* __restrict__ key, uint32_t* __restrict__ iota) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const int32_t* t = reinterpret_cast<const int32_t*>(raw + i * stride);
    attr[i] = flip(t[0]);
    key[i] = flip(t[1]);
    if (iota) iota[i] = (uint32_t)i;
}

// first index of sorted a[0, n) that is >= v (lower) or > v (upper)
__device__ __forceinline__ uint64_t bound(const uint32_t* a, uint64_t n, uint32_t v, bool upper) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (upper ? a[mid] <= v : a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// each R tuple's range of S tuples (sorted by key) with its key
__global__ void k_r_ranges(const uint32_t* __restrict__ rkey, uint64_t nR, const uint32_t* __restrict__ skey_sorted,
                           uint64_t nS, uint32_t* __restrict__ start, uint32_t* __restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= nR) return;
    const uint32_t k = rkey[i];
    const uint64_t lo = bound(skey_sorted, nS, k, false), hi = bound(skey_sorted, nS, k, true);
    start[i] = (uint32_t)lo;
    cnt[i] = (uint32_t)(hi - lo);
}

// the R tuple of path p: the last i with off[i] <= p
__device__ __forceinline__ uint32_t r_of_path(const uint64_t* off, uint64_t nR, uint64_t p) {
    uint64_t lo = 0, hi = nR;  // off[0] = 0 <= p
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (off[mid] <= p) lo = mid;
        else hi = mid;
    }
    return (uint32_t)lo;
}

// one (x, z) key per join path (all ones: dropped by exclude_equal)
__global__ void k_path_keys(const uint64_t* __restrict__ off, uint64_t nR, uint64_t P, const uint32_t* __restrict__ rattr,
                            const uint32_t* __restrict__ start, const uint32_t* __restrict__ sidx,
                            const uint32_t* __restrict__ sattr, uint32_t exclude_equal, uint64_t* __restrict__ key,
                            uint32_t* __restrict__ pid, unsigned long long* __restrict__ dropped) {
    const uint64_t p = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (p >= P) return;
    const uint32_t i = r_of_path(off, nR, p);
    const uint32_t j = sidx[start[i] + (uint32_t)(p - off[i])];
    const uint32_t x = rattr[i], z = sattr[j];
    const bool drop = exclude_equal && x == z;
    key[p] = drop ? ~0ull : ((uint64_t)x << 32) | z;
    pid[p] = (uint32_t)p;
    if (drop) atomicAdd(dropped, 1ull);
}

// group g: x, z, dense z key; path values in group order
__global__ void k_group_fields(const uint64_t* __restrict__ gkey, uint64_t G, const uint32_t* __restrict__ uz,
                               uint64_t nz, int32_t* __restrict__ gx, int32_t* __restrict__ gz,
                               uint32_t* __restrict__ gzkey) {
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= G) return;
    const uint64_t k = gkey[g];
    const uint32_t z = (uint32_t)k;
    gx[g] = unflip((uint32_t)(k >> 32));
    gz[g] = unflip(z);
    gzkey[g] = (uint32_t)bound(uz, nz, z, false);
}

__global__ void k_gather_values(const uint32_t* __restrict__ pid_sorted, uint64_t Q, const uint64_t* __restrict__ off,
                                uint64_t nR, const uint32_t* __restrict__ start, const uint32_t* __restrict__ sidx,
                                const uint8_t* __restrict__ rawR, const uint8_t* __restrict__ rawS, uint32_t stride,
                                uint32_t voff, uint32_t vsize, uint8_t* __restrict__ v1, uint8_t* __restrict__ v2) {
    const uint64_t q = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (q >= Q) return;
    const uint64_t p = pid_sorted[q];
    const uint32_t i = r_of_path(off, nR, p);
    const uint32_t j = sidx[start[i] + (uint32_t)(p - off[i])];
    const uint8_t* a = rawR + (uint64_t)i * stride + voff;
    const uint8_t* b = rawS + (uint64_t)j * stride + voff;
    if (vsize == 8) {
        reinterpret_cast<uint2*>(v1)[q] = *reinterpret_cast<const uint2*>(a);
        reinterpret_cast<uint2*>(v2)[q] = *reinterpret_cast<const uint2*>(b);
    } else {
        reinterpret_cast<uint32_t*>(v1)[q] = *reinterpret_cast<const uint32_t*>(a);
        reinterpret_cast<uint32_t*>(v2)[q] = *reinterpret_cast<const uint32_t*>(b);
    }
}

__global__ void k_u32_to_u64(const uint32_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// ---- classical join-project / COUNT (low OUT_J) ----
// sortable join key and attribute of each S tuple
__global__ void k_split_s(const int2* __restrict__ S, uint64_t n, uint32_t* __restrict__ key, uint32_t* __restrict__ attr) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const int2 t = S[i];
    key[i] = flip(t.y);
    attr[i] = flip(t.x);
}

__global__ void k_r_ranges2(const int2* __restrict__ R, uint64_t nR, const uint32_t* __restrict__ skey_sorted,
                            uint64_t nS, uint32_t* __restrict__ start, uint64_t* __restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= nR) return;
    const uint32_t k = flip(R[i].y);
    const uint64_t lo = bound(skey_sorted, nS, k, false), hi = bound(skey_sorted, nS, k, true);
    start[i] = (uint32_t)lo;
    cnt[i] = hi - lo;
}

// one (x, z) key per join path (all ones: dropped by exclude_equal)
__global__ void k_row_keys(const uint64_t* __restrict__ off, uint64_t nR, uint64_t P, const int2* __restrict__ R,
                           const uint32_t* __restrict__ start, const uint32_t* __restrict__ sattr_sorted,
                           uint32_t exclude_equal, uint64_t* __restrict__ key, unsigned long long* __restrict__ dropped) {
    const uint64_t p = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (p >= P) return;
    const uint32_t i = r_of_path(off, nR, p);
    const uint32_t x = flip(R[i].x), z = sattr_sorted[start[i] + (uint32_t)(p - off[i])];
    const bool drop = exclude_equal && x == z;
    key[p] = drop ? ~0ull : ((uint64_t)x << 32) | z;
    const uint64_t m = __ballot(drop);
    if ((threadIdx.x & 63u) == 0u && m) atomicAdd(dropped, (unsigned long long)__popcll(m));
}

// rows (materialised up to cap) and the fingerprint of the distinct keys
template <bool COUNT>
__global__ __launch_bounds__(256) void k_classical_rows(const uint64_t* __restrict__ gkey,
                                                        const uint32_t* __restrict__ gcnt, uint64_t G, void* out,
                                                        uint64_t cap, uint32_t fp, dim3h::Scalars* sc) {
    __shared__ uint64_t sh[5];
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint64_t s1 = 0, s2 = 0, cs = 0;
    if (g < G) {
        const uint64_t k = gkey[g];
        const int32_t x = unflip((uint32_t)(k >> 32)), z = unflip((uint32_t)k);
        const uint32_t c = gcnt[g];
        if (fp) {
            const uint64_t gz = dim3d::zfactor(z);
            s1 = dim3d::xfactor1(x) * gz;
            s2 = dim3d::xfactor2(x) * (COUNT ? (uint64_t)c : 1ull) * gz;
            cs = COUNT ? c : 0u;
        } else if (g < cap) {
            if (COUNT) {
                dim3_count_row r;
                r.x = x;
                r.z = z;
                r.count = (int64_t)c;
                reinterpret_cast<dim3_count_row*>(out)[g] = r;
            } else {
                reinterpret_cast<int2*>(out)[g] = make_int2(x, z);
            }
        }
    }
    if (!fp) return;
    s1 = dim3d::block_reduce_sum<256, uint64_t>(s1, sh);
    s2 = dim3d::block_reduce_sum<256, uint64_t>(s2, sh);
    cs = dim3d::block_reduce_sum<256, uint64_t>(cs, sh);
    if (threadIdx.x == 0) {
        atomicAdd((unsigned long long*)&sc->fp_sum, (unsigned long long)s1);
        atomicAdd((unsigned long long*)&sc->fp_sum2, (unsigned long long)s2);
        if (COUNT) atomicAdd((unsigned long long*)&sc->fp_csum, (unsigned long long)cs);
    }
}

// Device buffers of one call, freed in order on the stream.
struct Pool {
    hipStream_t st;
    void* p[48];
    int n = 0;
    explicit Pool(hipStream_t s) : st(s) {}
    template <typename T>
    hipError_t get(T** out, uint64_t count) {
        if (n == 48) return hipErrorOutOfMemory;
        void* q = nullptr;
        const hipError_t e = dim3h::dev_malloc_async(&q, std::max<uint64_t>(count, 1) * sizeof(T), st);
        if (e != hipSuccess) return e;
        p[n++] = q;
        *out = (T*)q;
        return hipSuccess;
    }
    ~Pool() {
        for (int i = n - 1; i >= 0; i--) dim3h::dev_free_async(p[i], st);
    }
};

}  // namespace

namespace dim3h {

#define PCHK(x)                                  \
    do {                                         \
        const hipError_t e_ = (x);               \
        if (e_ != hipSuccess) {                  \
            *err = #x;                           \
            *err += ": ";                        \
            *err += hipGetErrorString(e_);       \
            return DIM3_ERR_DEVICE;              \
        }                                        \
    } while (0)

int grouped_paths(const void* R, uint64_t nR, const void* S, uint64_t nS, uint32_t stride, uint32_t voff, uint32_t vsize,
                  bool exclude_equal, hipStream_t st, dim3_grouped* out, std::string* err) {
    Pool pool(st);
    uint8_t *rawR, *rawS;
    PCHK(pool.get(&rawR, std::max<uint64_t>(nR, 1) * stride));
    PCHK(pool.get(&rawS, std::max<uint64_t>(nS, 1) * stride));
    if (nR) PCHK(hipMemcpyAsync(rawR, R, nR * stride, hipMemcpyHostToDevice, st));
    if (nS) PCHK(hipMemcpyAsync(rawS, S, nS * stride, hipMemcpyHostToDevice, st));
    uint32_t *rattr, *rkey, *sattr, *skey, *sidx0, *skey_s, *sidx;
    PCHK(pool.get(&rattr, nR));
    PCHK(pool.get(&rkey, nR));
    PCHK(pool.get(&sattr, nS));
    PCHK(pool.get(&skey, nS));
    PCHK(pool.get(&sidx0, nS));
    PCHK(pool.get(&skey_s, nS));
    PCHK(pool.get(&sidx, nS));
    if (nR) k_split_raw<<<blocks_for(nR), 256, 0, st>>>(rawR, nR, stride, rattr, rkey, nullptr);
    if (nS) k_split_raw<<<blocks_for(nS), 256, 0, st>>>(rawS, nS, stride, sattr, skey, sidx0);
    PCHK(hipGetLastError());

    // 1. S tuple indices by join key
    size_t tb = 0, tb2 = 0;
    PCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, skey, skey_s, sidx0, sidx, nS, 0, 32, st));
    void* tmp = nullptr;
    uint8_t* tmp8;
    PCHK(pool.get(&tmp8, tb));
    tmp = tmp8;
    if (nS) PCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, skey, skey_s, sidx0, sidx, nS, 0, 32, st));

    // 2. S range of each R tuple, path offsets
    uint32_t *start, *cnt;
    uint64_t *cnt64, *off;
    PCHK(pool.get(&start, nR));
    PCHK(pool.get(&cnt, nR));
    PCHK(pool.get(&cnt64, nR + 1));
    PCHK(pool.get(&off, nR + 1));
    if (nR) k_r_ranges<<<blocks_for(nR), 256, 0, st>>>(rkey, nR, skey_s, nS, start, cnt);
    k_u32_to_u64<<<blocks_for(nR), 256, 0, st>>>(cnt, nR, cnt64);
    PCHK(hipMemsetAsync(cnt64 + nR, 0, sizeof(uint64_t), st));
    PCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, cnt64, off, nR + 1, st));
    uint8_t* tmp2;
    PCHK(pool.get(&tmp2, tb2));
    PCHK(hipcub::DeviceScan::ExclusiveSum(tmp2, tb2, cnt64, off, nR + 1, st));
    uint64_t P = 0;
    PCHK(copy_wait(&P, off + nR, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (P >= (1ull << 31)) {
        *err = "grouped join paths take fewer than 2^31 join paths (the query has " + std::to_string(P) + ")";
        return DIM3_ERR_UNSUPPORTED;
    }

    // 3. path keys, stable sort of path indices by (x, z)
    uint64_t *key, *key_s;
    uint32_t *pid, *pid_s;
    unsigned long long* dropped;
    PCHK(pool.get(&key, P));
    PCHK(pool.get(&key_s, P));
    PCHK(pool.get(&pid, P));
    PCHK(pool.get(&pid_s, P));
    PCHK(pool.get(&dropped, 1));
    PCHK(hipMemsetAsync(dropped, 0, sizeof(unsigned long long), st));
    if (P)
        k_path_keys<<<blocks_for(P), 256, 0, st>>>(off, nR, P, rattr, start, sidx, sattr, exclude_equal ? 1u : 0u,
                                                     key, pid, dropped);
    PCHK(hipGetLastError());
    size_t tb3 = 0;
    PCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, key, key_s, pid, pid_s, P, 0, 64, st));
    uint8_t* tmp3;
    PCHK(pool.get(&tmp3, tb3));
    if (P) PCHK(hipcub::DeviceRadixSort::SortPairs(tmp3, tb3, key, key_s, pid, pid_s, P, 0, 64, st));
    unsigned long long ndrop = 0;
    PCHK(copy_wait(&ndrop, dropped, sizeof(ndrop), hipMemcpyDeviceToHost, st));
    const uint64_t Q = P - ndrop;  // dropped paths sort last

    // 4. groups, dense z keys, values in group order
    uint64_t* gkey;
    uint32_t *gcnt, *ngroups_d;
    PCHK(pool.get(&gkey, Q));
    PCHK(pool.get(&gcnt, Q));
    PCHK(pool.get(&ngroups_d, 1));
    size_t tb4 = 0;
    PCHK(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb4, key_s, gkey, gcnt, ngroups_d, (int)Q, st));
    uint8_t* tmp4;
    PCHK(pool.get(&tmp4, tb4));
    PCHK(hipMemsetAsync(ngroups_d, 0, sizeof(uint32_t), st));
    if (Q) PCHK(hipcub::DeviceRunLengthEncode::Encode(tmp4, tb4, key_s, gkey, gcnt, ngroups_d, (int)Q, st));
    uint32_t G32 = 0;
    PCHK(copy_wait(&G32, ngroups_d, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    const uint64_t G = G32;

    // distinct S attribute values (sorted): the dense z keys
    uint32_t *sattr_s, *uz, *nz_d;
    PCHK(pool.get(&sattr_s, nS));
    PCHK(pool.get(&uz, nS));
    PCHK(pool.get(&nz_d, 1));
    size_t tb5 = 0, tb6 = 0;
    PCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb5, sattr, sattr_s, nS, 0, 32, st));
    PCHK(hipcub::DeviceSelect::Unique(nullptr, tb6, sattr_s, uz, nz_d, (int64_t)nS, st));
    uint8_t* tmp5;
    PCHK(pool.get(&tmp5, std::max(tb5, tb6)));
    PCHK(hipMemsetAsync(nz_d, 0, sizeof(uint32_t), st));
    if (nS) {
        PCHK(hipcub::DeviceRadixSort::SortKeys(tmp5, tb5, sattr, sattr_s, nS, 0, 32, st));
        PCHK(hipcub::DeviceSelect::Unique(tmp5, tb6, sattr_s, uz, nz_d, (int64_t)nS, st));
    }
    uint32_t nz = 0;
    PCHK(copy_wait(&nz, nz_d, sizeof(uint32_t), hipMemcpyDeviceToHost, st));

    int32_t *gx, *gz;
    uint32_t* gzkey;
    uint64_t *gcnt64, *gstart;
    uint8_t *v1, *v2;
    PCHK(pool.get(&gx, G));
    PCHK(pool.get(&gz, G));
    PCHK(pool.get(&gzkey, G));
    PCHK(pool.get(&gcnt64, G + 1));
    PCHK(pool.get(&gstart, G + 1));
    PCHK(pool.get(&v1, Q * vsize));
    PCHK(pool.get(&v2, Q * vsize));
    if (G) k_group_fields<<<blocks_for(G), 256, 0, st>>>(gkey, G, uz, nz, gx, gz, gzkey);
    k_u32_to_u64<<<blocks_for(G), 256, 0, st>>>(gcnt, G, gcnt64);
    PCHK(hipMemsetAsync(gcnt64 + G, 0, sizeof(uint64_t), st));
    size_t tb7 = 0;
    PCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb7, gcnt64, gstart, G + 1, st));
    uint8_t* tmp7;
    PCHK(pool.get(&tmp7, tb7));
    PCHK(hipcub::DeviceScan::ExclusiveSum(tmp7, tb7, gcnt64, gstart, G + 1, st));
    if (Q)
        k_gather_values<<<blocks_for(Q), 256, 0, st>>>(pid_s, Q, off, nR, start, sidx, rawR, rawS, stride, voff, vsize,
                                                       v1, v2);
    PCHK(hipGetLastError());

    // to the host
    out->n_groups = G;
    out->n_paths = Q;
    out->nz = nz;
    out->value_size = vsize;
    out->x = (int32_t*)malloc(std::max<uint64_t>(G, 1) * sizeof(int32_t));
    out->z = (int32_t*)malloc(std::max<uint64_t>(G, 1) * sizeof(int32_t));
    out->zkey = (uint32_t*)malloc(std::max<uint64_t>(G, 1) * sizeof(uint32_t));
    out->path_start = (uint64_t*)malloc((G + 1) * sizeof(uint64_t));
    out->v1 = malloc(std::max<uint64_t>(Q, 1) * vsize);
    out->v2 = malloc(std::max<uint64_t>(Q, 1) * vsize);
    if (!out->x || !out->z || !out->zkey || !out->path_start || !out->v1 || !out->v2) {
        *err = "host allocation of the grouped paths failed";
        return DIM3_ERR_OOM;
    }
    if (G) {
        PCHK(hipMemcpyAsync(out->x, gx, G * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        PCHK(hipMemcpyAsync(out->z, gz, G * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        PCHK(hipMemcpyAsync(out->zkey, gzkey, G * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    }
    PCHK(hipMemcpyAsync(out->path_start, gstart, (G + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (Q) {
        PCHK(hipMemcpyAsync(out->v1, v1, Q * vsize, hipMemcpyDeviceToHost, st));
        PCHK(hipMemcpyAsync(out->v2, v2, Q * vsize, hipMemcpyDeviceToHost, st));
    }
    PCHK(hipStreamSynchronize(st));
    return DIM3_OK;
}

// Classical join-project / COUNT of device relations (after the swap rule):
// the counterpart of the reference's Radix_hash_join_and_project
// (dim3.hpp:273-391), chosen by f1 for low OUT_J.  S sorted by join key
// with its attribute, each R tuple's S range by binary search, one (x, z)
// key per join path, a radix sort of the keys and a run-length encoding: the
// runs are the rows, their lengths the COUNTs.  Rows in (x, z) order.
// Fills sc->out_n (rows), sc->paths (join paths) and, fingerprint mode, the
// fingerprint sums; otherwise get_out(rows) names the output buffer and its
// capacity once the row count is known, and min(rows, capacity) rows are
// written.
int classical_join(const int2* R, uint64_t nR, const int2* S, uint64_t nS, bool count_mode, bool exclude_equal,
                   bool fp, const std::function<int(uint64_t, void**, uint64_t*)>& get_out, Scalars* sc,
                   hipStream_t st, std::string* err) {
    Pool pool(st);
    uint32_t *skey, *sattr, *skey_s, *sattr_s, *start;
    uint64_t *cnt, *off;
    PCHK(pool.get(&skey, nS));
    PCHK(pool.get(&sattr, nS));
    PCHK(pool.get(&skey_s, nS));
    PCHK(pool.get(&sattr_s, nS));
    PCHK(pool.get(&start, nR));
    PCHK(pool.get(&cnt, nR + 1));
    PCHK(pool.get(&off, nR + 1));
    k_split_s<<<blocks_for(nS), 256, 0, st>>>(S, nS, skey, sattr);
    size_t tb = 0, tb2 = 0;
    PCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, skey, skey_s, sattr, sattr_s, nS, 0, 32, st));
    PCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, cnt, off, nR + 1, st));
    uint8_t* tmp;
    PCHK(pool.get(&tmp, std::max(tb, tb2)));
    PCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, skey, skey_s, sattr, sattr_s, nS, 0, 32, st));
    k_r_ranges2<<<blocks_for(nR), 256, 0, st>>>(R, nR, skey_s, nS, start, cnt);
    PCHK(hipMemsetAsync(cnt + nR, 0, sizeof(uint64_t), st));
    PCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, cnt, off, nR + 1, st));
    uint64_t P = 0;
    PCHK(copy_wait(&P, off + nR, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (P >= (1ull << 31)) {
        *err = "the classical join takes fewer than 2^31 join paths (the query has " + std::to_string(P) + ")";
        return DIM3_ERR_UNSUPPORTED;
    }
    uint64_t *key, *key_s, *gkey;
    uint32_t *gcnt, *nruns;
    unsigned long long* dropped;
    PCHK(pool.get(&key, P));
    PCHK(pool.get(&key_s, P));
    PCHK(pool.get(&gkey, P));
    PCHK(pool.get(&gcnt, P));
    PCHK(pool.get(&nruns, 1));
    PCHK(pool.get(&dropped, 1));
    PCHK(hipMemsetAsync(dropped, 0, sizeof(unsigned long long), st));
    PCHK(hipMemsetAsync(nruns, 0, sizeof(uint32_t), st));
    if (P) k_row_keys<<<blocks_for(P), 256, 0, st>>>(off, nR, P, R, start, sattr_s, exclude_equal ? 1u : 0u, key, dropped);
    size_t tb3 = 0, tb4 = 0;
    PCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb3, key, key_s, P, 0, 64, st));
    PCHK(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb4, key_s, gkey, gcnt, nruns, (int)P, st));
    uint8_t* tmp3;
    PCHK(pool.get(&tmp3, std::max(tb3, tb4)));
    if (P) {
        PCHK(hipcub::DeviceRadixSort::SortKeys(tmp3, tb3, key, key_s, P, 0, 64, st));
        PCHK(hipcub::DeviceRunLengthEncode::Encode(tmp3, tb4, key_s, gkey, gcnt, nruns, (int)P, st));
    }
    uint32_t hr[2] = {0, 0};
    unsigned long long hd = 0;
    PCHK(hipMemcpyAsync(&hr[0], nruns, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PCHK(copy_wait(&hd, dropped, sizeof(hd), hipMemcpyDeviceToHost, st));
    const uint64_t G = hr[0] - (hd ? 1u : 0u);  // the dropped paths' key sorts last
    void* dout = nullptr;
    uint64_t dcap = 0;
    if (!fp) {
        const int rc = get_out(G, &dout, &dcap);
        if (rc != DIM3_OK) {
            *err = "no output buffer for " + std::to_string(G) + " rows";
            return rc;
        }
    }
    if (G && count_mode)
        k_classical_rows<true><<<blocks_for(G), 256, 0, st>>>(gkey, gcnt, G, dout, dcap, fp ? 1u : 0u, sc);
    else if (G)
        k_classical_rows<false><<<blocks_for(G), 256, 0, st>>>(gkey, gcnt, G, dout, dcap, fp ? 1u : 0u, sc);
    PCHK(hipGetLastError());
    const unsigned long long tot[2] = {(unsigned long long)G, (unsigned long long)P};
    PCHK(hipMemcpyAsync(&sc->out_n, &tot[0], sizeof(unsigned long long), hipMemcpyHostToDevice, st));
    PCHK(hipMemcpyAsync(&sc->paths, &tot[1], sizeof(unsigned long long), hipMemcpyHostToDevice, st));
    PCHK(hipStreamSynchronize(st));
    return DIM3_OK;
}

#undef PCHK

}  // namespace dim3h
