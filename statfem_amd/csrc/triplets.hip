// Triplet::finishAssembly + setCompressedRowArrays on the GPU
// (corlib/Triplet.hpp:126-170,264-339): element contributions given as
// (row, col, value) triplets in insertion order become the CSR operator with
// rows in order, columns ascending, equal keys merged by summation and
// explicitly inserted zeros kept. This is the path of callers that insert
// per-element matrices through the SystemSolve concept
// (corlib/SystemSolveEigenSparse.hpp:171-182) instead of uploading a mesh.
//
// Keys (row << 32 | col) are sorted with a stable LSD radix sort, so entries
// with equal keys stay in insertion order and their sum is formed in that
// order by one thread per merged entry (deterministic). The sorted-to-original
// permutation and the merged-entry starts are kept, so the values of a later
// assembly with the same insertion sequence are merged without sorting again.
#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace sfh {
namespace {

__global__ void k_make_keys(long long n, const int* __restrict__ rows, const int* __restrict__ cols,
                            unsigned long long* __restrict__ keys, int* __restrict__ idx) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ((unsigned long long)(unsigned)rows[i] << 32) | (unsigned)cols[i];
    idx[i] = (int)i;
}

__global__ void k_key_heads(long long n, const unsigned long long* __restrict__ keys, int* __restrict__ head) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// merged entry u = uid[i] starts at sorted position i: column, start, row count
__global__ void k_unique(long long n, const unsigned long long* __restrict__ keys, const int* __restrict__ head,
                         const int* __restrict__ uid, int* __restrict__ colind, long long* __restrict__ start,
                         int* __restrict__ rowcnt) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    const int u = uid[i];
    colind[u] = (int)(keys[i] & 0xffffffffull);
    start[u] = i;
    atomicAdd(&rowcnt[(int)(keys[i] >> 32)], 1);
}

// vals[u] = sum over sorted positions start[u] .. start[u+1]-1 of v[perm[.]], in order
__global__ void k_merge_values(long long nnz, const long long* __restrict__ start, const int* __restrict__ perm,
                               const double* __restrict__ v, double* __restrict__ vals, float* __restrict__ vals32) {
    const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nnz) return;
    double s = 0.0;
    for (long long p = start[u]; p < start[u + 1]; ++p) s += v[perm[p]];
    vals[u] = s;
    vals32[u] = (float)s;
}

__global__ void k_row_diag(int n, const int* __restrict__ rp, const int* __restrict__ ci, int* __restrict__ diag,
                           int* __restrict__ bad) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int pos = -1;
    for (int p = rp[r]; p < rp[r + 1]; ++p)
        if (ci[p] == r) pos = p;
    diag[r] = pos;
    if (pos < 0) atomicMin(bad, r);
}

inline unsigned nbk(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int triplets_build(sfh_system* s, long long n, const int* rows, const int* cols, const double* vals) {
    hipStream_t st = s->ctx->stream;
    const int N = s->ndof;
    DevBuf dr, dc, keys_in, keys_out, idx_in, head, uid, tmp, cnt, bad;
    SFH_TRY(dr.ensure(sizeof(int) * n));
    SFH_TRY(dc.ensure(sizeof(int) * n));
    SFH_CHECK_HIP(hipMemcpyAsync(dr.p, rows, sizeof(int) * n, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(dc.p, cols, sizeof(int) * n, hipMemcpyHostToDevice, st));
    SFH_TRY(keys_in.ensure(8 * n));
    SFH_TRY(keys_out.ensure(8 * n));
    SFH_TRY(idx_in.ensure(sizeof(int) * n));
    SFH_TRY(s->coo_perm.ensure(sizeof(int) * n));
    k_make_keys<<<nbk(n), 256, 0, st>>>(n, dr.as<int>(), dc.as<int>(), keys_in.as<unsigned long long>(),
                                        idx_in.as<int>());
    SFH_LAUNCH_CHECK();
    int rbits = 1;
    while ((1LL << rbits) < N) ++rbits;
    size_t tb = 0;
    SFH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys_in.as<unsigned long long>(),
                                                     keys_out.as<unsigned long long>(), idx_in.as<int>(),
                                                     s->coo_perm.as<int>(), (int)n, 0, 32 + rbits, st));
    SFH_TRY(tmp.ensure(tb + 16));
    SFH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys_in.as<unsigned long long>(),
                                                     keys_out.as<unsigned long long>(), idx_in.as<int>(),
                                                     s->coo_perm.as<int>(), (int)n, 0, 32 + rbits, st));
    SFH_TRY(head.ensure(sizeof(int) * n));
    SFH_TRY(uid.ensure(sizeof(int) * n));
    k_key_heads<<<nbk(n), 256, 0, st>>>(n, keys_out.as<unsigned long long>(), head.as<int>());
    SFH_LAUNCH_CHECK();
    size_t tb2 = 0;
    SFH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, head.as<int>(), uid.as<int>(), (int)n, st));
    DevBuf tmp2;
    SFH_TRY(tmp2.ensure(tb2 + 16));
    SFH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp2.p, tb2, head.as<int>(), uid.as<int>(), (int)n, st));
    int last_uid = 0, last_head = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&last_uid, uid.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(&last_head, head.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    const long long nnz = (long long)last_uid + last_head;
    s->nnz = nnz;
    s->ncoo = n;
    SFH_TRY(s->colind.ensure(sizeof(int) * nnz));
    SFH_TRY(s->coo_start.ensure(sizeof(long long) * (nnz + 1)));
    SFH_TRY(cnt.ensure(sizeof(int) * (N + 1)));
    SFH_CHECK_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * (N + 1), st));
    k_unique<<<nbk(n), 256, 0, st>>>(n, keys_out.as<unsigned long long>(), head.as<int>(), uid.as<int>(),
                                     s->colind.as<int>(), s->coo_start.as<long long>(), cnt.as<int>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemcpyAsync(s->coo_start.as<long long>() + nnz, &n, sizeof(long long), hipMemcpyHostToDevice,
                                 st));
    SFH_TRY(s->rowptr.ensure(sizeof(int) * (N + 1)));
    size_t tb3 = 0;
    SFH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, cnt.as<int>(), s->rowptr.as<int>(), N + 1, st));
    DevBuf tmp3;
    SFH_TRY(tmp3.ensure(tb3 + 16));
    SFH_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp3.p, tb3, cnt.as<int>(), s->rowptr.as<int>(), N + 1, st));
    SFH_TRY(s->diagpos.ensure(sizeof(int) * N));
    SFH_TRY(bad.ensure(sizeof(int)));
    const int big = 0x7fffffff;
    SFH_CHECK_HIP(hipMemcpyAsync(bad.p, &big, sizeof(int), hipMemcpyHostToDevice, st));
    k_row_diag<<<nbk(N), 256, 0, st>>>(N, s->rowptr.as<int>(), s->colind.as<int>(), s->diagpos.as<int>(),
                                       bad.as<int>());
    SFH_LAUNCH_CHECK();
    std::vector<int> rp(N + 1);
    int hbad = big;
    SFH_CHECK_HIP(hipMemcpyAsync(rp.data(), s->rowptr.p, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (hbad != big) {
        set_error("row %d has no diagonal entry", hbad);
        return SFH_EARG;
    }
    int mx = 0;
    for (int r = 0; r < N; ++r) mx = std::max(mx, rp[r + 1] - rp[r]);
    s->maxrow = mx;
    SFH_TRY(s->vals.ensure(sizeof(double) * nnz));
    SFH_TRY(s->vals32.ensure(sizeof(float) * nnz));
    SFH_TRY(s->F.ensure(sizeof(double) * N));
    SFH_TRY(s->cmask.ensure(N));
    SFH_TRY(s->dinv32.ensure(sizeof(float) * N));
    SFH_TRY(s->dinv64.ensure(sizeof(double) * N));
    SFH_TRY(s->diag32.ensure(sizeof(float) * N));
    SFH_TRY(s->diag64.ensure(sizeof(double) * N));
    SFH_TRY(s->cv32.ensure(sizeof(int2) * nnz));
    SFH_CHECK_HIP(hipMemsetAsync(s->F.p, 0, sizeof(double) * N, st));
    SFH_CHECK_HIP(hipMemsetAsync(s->cmask.p, 0, N, st));
    return triplets_values(s, n, vals);
}

int triplets_values(sfh_system* s, long long n, const double* vals) {
    if (n != s->ncoo) { set_error("expected %lld triplet values, got %lld", s->ncoo, n); return SFH_EARG; }
    hipStream_t st = s->ctx->stream;
    DevBuf dv;
    SFH_TRY(dv.ensure(sizeof(double) * n));
    SFH_CHECK_HIP(hipMemcpyAsync(dv.p, vals, sizeof(double) * n, hipMemcpyHostToDevice, st));
    k_merge_values<<<nbk(s->nnz), 256, 0, st>>>(s->nnz, s->coo_start.as<long long>(), s->coo_perm.as<int>(),
                                                dv.as<double>(), s->vals.as<double>(), s->vals32.as<float>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemsetAsync(s->cmask.p, 0, s->ndof, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    s->assembled = true;
    s->dirty = true;
    s->nc = 0;
    return SFH_OK;
}

}  // namespace sfh

using namespace sfh;

extern "C" {

int sfh_system_create_triplets(sfh_ctx* c, int ndof, long long n, const int* rows, const int* cols,
                               const double* vals, sfh_system** out) {
    if (!c || !out || ndof <= 0 || n <= 0 || !rows || !cols || !vals) { set_error("bad argument"); return SFH_EARG; }
    *out = nullptr;
    if (n >= (1LL << 31)) { set_error("too many triplets"); return SFH_EARG; }
    for (long long i = 0; i < n; ++i)
        if (rows[i] < 0 || rows[i] >= ndof || cols[i] < 0 || cols[i] >= ndof) {
            set_error("triplet %lld (%d, %d) out of range", i, rows[i], cols[i]);
            return SFH_EARG;
        }
    SFH_CHECK_HIP(hipSetDevice(c->device));
    sfh_system* s = new sfh_system;
    s->ctx = c;
    s->kind = 0;
    s->dofpn = 1;
    s->nnodes = ndof;
    s->ndof = ndof;
    const int rc = sfh::triplets_build(s, n, rows, cols, vals);
    if (rc != SFH_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return SFH_OK;
}

int sfh_system_set_triplet_values(sfh_system* s, long long n, const double* vals) {
    if (!s || !vals) { set_error("bad argument"); return SFH_EARG; }
    if (!s->ncoo) { set_error("system was not built from triplets"); return SFH_ESTATE; }
    return sfh::triplets_values(s, n, vals);
}

int sfh_system_set_rhs(sfh_system* s, const double* F) {
    if (!s || !F) { set_error("bad argument"); return SFH_EARG; }
    SFH_CHECK_HIP(hipMemcpyAsync(s->F.p, F, sizeof(double) * s->ndof, hipMemcpyHostToDevice, s->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SFH_OK;
}

}  // extern "C"
