// extern "C" entry points of libstatfem_hip (declared in include/statfem_hip.h).
// No exception or abort crosses this boundary: every call returns a status.
#include <algorithm>
#include <cstdio>
#include <cstring>

#include <string>
#include <thread>

#include "common.hpp"

namespace sfh {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int DevBuf::ensure(size_t b) {
    if (b <= bytes && p) return SFH_OK;
    release();
    if (b == 0) b = 8;
    if (hipMalloc(&p, b) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        bytes = 0;
        set_error("device allocation of %zu bytes failed", b);
        return SFH_ENOMEM;
    }
    bytes = b;
    return SFH_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int PinnedBuf::ensure(size_t b) {
    if (b <= bytes && p) return SFH_OK;
    release();
    if (b == 0) b = 8;
    if (hipHostMalloc(&p, b, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        bytes = 0;
        set_error("page-locked host allocation of %zu bytes failed", b);
        return SFH_ENOMEM;
    }
    bytes = b;
    return SFH_OK;
}

void PinnedBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

namespace {

inline unsigned nb(long long n, int b) { return (unsigned)((n + b - 1) / b); }

__global__ void k_scatter_pt(int ns, int pn, const int* cols, const double* vals, int ld, double* Pt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    for (int q = 0; q < pn; ++q) Pt[(size_t)cols[s * pn + q] * ld + s] += vals[s * pn + q];
}

// y += a P^T on the entries P^T has (the sparse sensor rows as (col, val)
// pairs): the entries of P^T are read from its dense copy, so every touched
// entry gets the value a dense axpy would add; a DOF repeated in one row is
// added once
__global__ void k_pt_axpy(int ns, int pn, const int* cols, double a, const double* Pt, int ld, double* y) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    for (int q = 0; q < pn; ++q) {
        const int c = cols[s * pn + q];
        bool seen = false;
        for (int u = 0; u < q; ++u) seen = seen || cols[s * pn + u] == c;
        if (!seen) y[(size_t)c * ld + s] += a * Pt[(size_t)c * ld + s];
    }
}

__global__ void k_mask_zero(int n, const unsigned char* m, double* v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && m[i]) v[i] = 0.0;
}

__global__ void k_dof_gather(int n, int k, int stride, int comp, const double* __restrict__ w,
                             const double* __restrict__ src, double* __restrict__ dst) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * k) return;
    const int j = (int)(t / k), c = (int)(t % k);
    dst[t] = w[j] * src[((size_t)j * stride + comp) * k + c];
}

__global__ void k_dof_scatter(int n, int k, int stride, int comp, const double* __restrict__ w,
                              const double* __restrict__ src, double* __restrict__ dst) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * stride * k) return;
    const long long row = t / k;
    const int c = (int)(t % k), j = (int)(row / stride), q = (int)(row % stride);
    dst[t] = q == comp ? w[j] * src[(size_t)j * k + c] : 0.0;
}

__global__ void k_rows_scale(long long nk, int k, const double* __restrict__ w, const double* __restrict__ X,
                             double* __restrict__ Y) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nk) Y[t] = w[t / k] * X[t];
}

__global__ void k_axpy(long long n, double a, const double* x, double* y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// S = Cu + K_s / (rho^2 no) and B = [Cu | ybar/rho - mean] (ns x (ns+1))
__global__ void k_post_build(int ns, int no, int dim, const double* xs, const double* Cu, const double* mean,
                             const double* Y, double eps, double rho, double sd, double ld, double* S, double* B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns * ns) return;
    const int i = t / ns, j = t % ns;
    double sq = 0.0;
    for (int d = 0; d < dim; ++d) {
        const double dd = xs[(size_t)i * dim + d] - xs[(size_t)j * dim + d];
        sq += dd * dd;
    }
    double k = sd * sd * exp(-sq / (2. * ld * ld));
    if (i == j) k += eps * eps;
    S[t] = Cu[t] + k / (rho * rho * no);
    B[(size_t)i * (ns + 1) + j] = Cu[t];
    if (j == 0) {
        double ys = 0.0;
        for (int q = 0; q < no; ++q) ys += Y[(size_t)q * ns + i];
        B[(size_t)i * (ns + 1) + ns] = ys / no / rho - mean[i];
    }
}

// B = I (n x n)
__global__ void k_eye(int n, double* B) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long long)n * n) B[t] = (t / n == t % n) ? 1.0 : 0.0;
}

// out_i = prior_i - sum_s H[i][s] F[i][s] (N rows of ns): the diagonal of
// F S^-1 F^T subtracted from the prior variance, one wave per row
__global__ __launch_bounds__(256) void k_post_var(int N, int ns, const double* __restrict__ F,
                                                  const double* __restrict__ H, const double* __restrict__ prior,
                                                  double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= N) return;
    const double* f = F + (size_t)row * ns;
    const double* h = H + (size_t)row * ns;
    double s = 0.0;
    for (int c = lane; c < ns; c += 64) s += f[c] * h[c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[row] = prior[row] - s;
}

// At (C x R) = A^T for A R x C, through 32 x 32 LDS tiles
__global__ __launch_bounds__(256) void k_transpose(int R, int C, const double* __restrict__ A, double* __restrict__ At) {
    __shared__ double t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8)
        if (r0 + r < R && c0 + tx < C) t[r][tx] = A[(size_t)(r0 + r) * C + c0 + tx];
    __syncthreads();
    for (int c = ty; c < 32; c += 8)
        if (c0 + c < C && r0 + tx < R) At[(size_t)(c0 + c) * R + r0 + tx] = t[tx][c];
}

// out = prior - T, elementwise (out may alias prior)
__global__ void k_sub(long long n, const double* prior, const double* T, double* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = prior[i] - T[i];
}

// cov = Cu - T[:, :ns], mean = mean + T[:, ns]
__global__ void k_post_finish(int ns, const double* Cu, const double* mean, const double* T, double* cov,
                              double* mpost) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns * ns) return;
    const int i = t / ns, j = t % ns;
    cov[t] = Cu[t] - T[(size_t)i * (ns + 1) + j];
    if (j == 0) mpost[i] = mean[i] + T[(size_t)i * (ns + 1) + ns];
}

}  // namespace
}  // namespace sfh

using namespace sfh;

extern "C" {

const char* sfh_status_string(int s) {
    switch (s) {
        case SFH_OK: return "ok";
        case SFH_EARG: return "bad argument";
        case SFH_ENEGJAC: return "negative Jacobian";
        case SFH_ESINGULAR: return "singular system";
        case SFH_ENOTSPD: return "covariance not positive definite";
        case SFH_EQUADFORM: return "non-negative quadratic form";
        case SFH_ENOCONV: return "solver did not converge";
        case SFH_EDEVICE: return "device error";
        case SFH_ENOMEM: return "out of device memory";
        case SFH_ESTATE: return "call out of order";
    }
    return "unknown status";
}

const char* sfh_last_error(void) { return g_err; }
const char* sfh_version(void) { return "statfem_hip 0.1 gfx950"; }

int sfh_ctx_create(int device, sfh_ctx** out) {
    if (!out) { set_error("null out"); return SFH_EARG; }
    *out = nullptr;
    int n = 0;
    SFH_CHECK_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) { set_error("device %d of %d", device, n); return SFH_EARG; }
    SFH_CHECK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SFH_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    double* scr = nullptr;
    SFH_TRY(gp_scratch_init(device, &scr));   // this GPU's fixed GP scratch (sfh_gp_noalloc)
    sfh_ctx* c = new sfh_ctx;
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    std::snprintf(c->arch, sizeof c->arch, "%s", prop.gcnArchName);
    c->gp_scratch = scr;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->tev0) != hipSuccess || hipEventCreate(&c->tev1) != hipSuccess) {
        set_error("stream/event creation failed");
        delete c;
        return SFH_EDEVICE;
    }
    *out = c;
    return SFH_OK;
}

int sfh_ctx_info(sfh_ctx* c, int* device, int* num_cu, char* arch, int len) {
    if (!c) { set_error("null context"); return SFH_EARG; }
    if (device) *device = c->device;
    if (num_cu) *num_cu = c->num_cu;
    if (arch && len > 0) std::snprintf(arch, (size_t)len, "%s", c->arch);
    return SFH_OK;
}

int sfh_ctx_destroy(sfh_ctx* c) {
    if (!c) return SFH_OK;
    (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) (void)hipFree(p);
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipEventDestroy(c->tev0);
    (void)hipEventDestroy(c->tev1);
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    if (c->aev0) (void)hipEventDestroy(c->aev0);
    if (c->aev1) (void)hipEventDestroy(c->aev1);
    (void)hipStreamDestroy(c->stream);
    (void)hipStreamDestroy(c->stream2);
    delete c;
    return SFH_OK;
}

int sfh_ctx_sync(sfh_ctx* c) {
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SFH_OK;
}

int sfh_malloc(sfh_ctx* c, size_t bytes, void** d) {
    if (!c || !d) { set_error("null argument"); return SFH_EARG; }
    if (hipMalloc(d, bytes ? bytes : 8) != hipSuccess) {
        (void)hipGetLastError();
        set_error("device allocation of %zu bytes failed", bytes);
        return SFH_ENOMEM;
    }
    c->allocs.push_back(*d);
    return SFH_OK;
}

int sfh_free(sfh_ctx* c, void* d) {
    if (!c || !d) return SFH_OK;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), d);
    if (it == c->allocs.end()) { set_error("pointer not allocated by sfh_malloc"); return SFH_EARG; }
    c->allocs.erase(it);
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    SFH_CHECK_HIP(hipFree(d));
    return SFH_OK;
}

int sfh_memcpy_h2d(sfh_ctx* c, void* dst, const void* src, size_t bytes) {
    SFH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SFH_OK;
}

int sfh_memcpy_d2h(sfh_ctx* c, void* dst, const void* src, size_t bytes) {
    SFH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SFH_OK;
}

int sfh_memcpy_d2d(sfh_ctx* c, void* dst, const void* src, size_t bytes) {
    SFH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return SFH_OK;
}

int sfh_memset(sfh_ctx* c, void* dst, int v, size_t bytes) {
    SFH_CHECK_HIP(hipMemsetAsync(dst, v, bytes, c->stream));
    return SFH_OK;
}

int sfh_dof_gather(sfh_ctx* c, int n, int k, int stride, int comp, const double* w, const double* src, double* dst) {
    if (!c || n <= 0 || k <= 0 || stride <= 0 || comp < 0 || comp >= stride || !w || !src || !dst) {
        set_error("bad argument");
        return SFH_EARG;
    }
    k_dof_gather<<<nb((long long)n * k, 256), 256, 0, c->stream>>>(n, k, stride, comp, w, src, dst);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int sfh_dof_scatter(sfh_ctx* c, int n, int k, int stride, int comp, const double* w, const double* src, double* dst) {
    if (!c || n <= 0 || k <= 0 || stride <= 0 || comp < 0 || comp >= stride || !w || !src || !dst) {
        set_error("bad argument");
        return SFH_EARG;
    }
    k_dof_scatter<<<nb((long long)n * stride * k, 256), 256, 0, c->stream>>>(n, k, stride, comp, w, src, dst);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int sfh_row_scale(sfh_ctx* c, int n, int k, const double* w, const double* X, double* Y) {
    if (!c || n <= 0 || k <= 0 || !w || !X || !Y) { set_error("bad argument"); return SFH_EARG; }
    k_rows_scale<<<nb((long long)n * k, 256), 256, 0, c->stream>>>((long long)n * k, k, w, X, Y);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int sfh_timer_start(sfh_ctx* c) {
    SFH_CHECK_HIP(hipEventRecord(c->tev0, c->stream));
    return SFH_OK;
}

int sfh_timer_stop(sfh_ctx* c, double* ms) {
    SFH_CHECK_HIP(hipEventRecord(c->tev1, c->stream));
    SFH_CHECK_HIP(hipEventSynchronize(c->tev1));
    float f = 0.f;
    SFH_CHECK_HIP(hipEventElapsedTime(&f, c->tev0, c->tev1));
    if (ms) *ms = f;
    return SFH_OK;
}

// ------------------------------------------------------------- system
int sfh_system_create(sfh_ctx* c, int kind, int dofpn, int nnodes, const double* coords, int nel,
                      const int* conn, sfh_system** out) {
    if (!c || !out || !coords || !conn) { set_error("null argument"); return SFH_EARG; }
    *out = nullptr;
    const int d = kind_dim(kind), nen = kind_nen(kind);
    if (!d) { set_error("unknown element kind %d", kind); return SFH_EARG; }
    if (dofpn != 1 && dofpn != 3) { set_error("dofpn must be 1 or 3"); return SFH_EARG; }
    if (nnodes <= 0 || nel <= 0) { set_error("empty mesh"); return SFH_EARG; }
    if (nel >= (1 << 28)) { set_error("too many elements"); return SFH_EARG; }
    for (long long i = 0; i < (long long)nel * nen; ++i)
        if (conn[i] < 0 || conn[i] >= nnodes) { set_error("connectivity entry %lld out of range", i); return SFH_EARG; }
    SFH_CHECK_HIP(hipSetDevice(c->device));
    sfh_system* s = new sfh_system;
    s->ctx = c;
    s->kind = kind;
    s->dofpn = dofpn;
    s->nnodes = nnodes;
    s->nel = nel;
    s->nen = nen;
    s->dim = d;
    int rc = s->coords.ensure(sizeof(double) * nnodes * d);
    if (rc == SFH_OK) rc = s->conn.ensure(sizeof(int) * (size_t)nel * nen);
    if (rc == SFH_OK &&
        (hipMemcpyAsync(s->coords.p, coords, sizeof(double) * nnodes * d, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
         hipMemcpyAsync(s->conn.p, conn, sizeof(int) * (size_t)nel * nen, hipMemcpyHostToDevice, c->stream) != hipSuccess)) {
        set_error("upload failed");
        rc = SFH_EDEVICE;
    }
    if (rc == SFH_OK) rc = build_pattern(s);
    if (rc != SFH_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return SFH_OK;
}

int sfh_system_destroy(sfh_system* s) {
    if (!s) return SFH_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    return SFH_OK;
}

int sfh_system_info(sfh_system* s, int* ndof, long long* nnz) {
    if (!s) { set_error("null system"); return SFH_EARG; }
    if (ndof) *ndof = s->ndof;
    if (nnz) *nnz = s->nnz;
    return SFH_OK;
}

int sfh_system_pattern(sfh_system* s, int* rowptr, int* colind) {
    if (!s) { set_error("null system"); return SFH_EARG; }
    if (rowptr)
        SFH_CHECK_HIP(hipMemcpyAsync(rowptr, s->rowptr.p, sizeof(int) * (s->ndof + 1), hipMemcpyDeviceToHost, s->ctx->stream));
    if (colind)
        SFH_CHECK_HIP(hipMemcpyAsync(colind, s->colind.p, sizeof(int) * (size_t)s->nnz, hipMemcpyDeviceToHost, s->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SFH_OK;
}

namespace {
// the lowest-numbered node that is not at its grid point (node i at
// (i mod nx, (i / nx) mod ny, i / (nx ny)) of the axes, to 1e-12), into *bad
__global__ void k_grid_nodes(int nnodes, int d, int nx, int ny, const double* __restrict__ X,
                             const double* __restrict__ gx, const double* __restrict__ gy,
                             const double* __restrict__ gz, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    const int idx[3] = {i % nx, (i / nx) % ny, i / (nx * ny)};
    const double* g[3] = {gx, gy, gz};
    for (int a = 0; a < d; ++a)
        if (!(fabs(X[(size_t)i * d + a] - g[a][idx[a]]) <= 1e-12)) {
            atomicMin(bad, i);
            return;
        }
}
}  // namespace

int sfh_system_set_grid(sfh_system* s, int nx, int ny, int nz, const double* gx, const double* gy, const double* gz) {
    if (s && s->kind == 0) { set_error("this system was built from triplets and has no mesh"); return SFH_ESTATE; }
    if (!s) { set_error("null system"); return SFH_EARG; }
    const int d = s->dim;
    const int n[3] = {nx, d > 1 ? ny : 1, d > 2 ? nz : 1};
    const double* g[3] = {gx, gy, gz};
    if ((long long)n[0] * n[1] * n[2] != s->nnodes) { set_error("grid size does not match node count"); return SFH_EARG; }
    // every node at its grid point, checked on the device against the axes
    DevBuf tmp;
    const size_t nax = (size_t)n[0] + (d > 1 ? n[1] : 0) + (d > 2 ? n[2] : 0);
    SFH_TRY(tmp.ensure(sizeof(double) * nax + sizeof(int)));
    double* ax = tmp.as<double>();
    int* bad = reinterpret_cast<int*>(ax + nax);
    for (int a = 0, o = 0; a < d; o += n[a], ++a)
        SFH_CHECK_HIP(hipMemcpyAsync(ax + o, g[a], sizeof(double) * n[a], hipMemcpyHostToDevice, s->ctx->stream));
    const int big = INT32_MAX;
    SFH_CHECK_HIP(hipMemcpyAsync(bad, &big, sizeof(int), hipMemcpyHostToDevice, s->ctx->stream));
    k_grid_nodes<<<(unsigned)((s->nnodes + 255) / 256), 256, 0, s->ctx->stream>>>(
        s->nnodes, d, n[0], n[1], s->coords.as<double>(), ax, d > 1 ? ax + n[0] : nullptr,
        d > 2 ? ax + n[0] + n[1] : nullptr, bad);
    SFH_LAUNCH_CHECK();
    int first = big;
    SFH_CHECK_HIP(hipMemcpyAsync(&first, bad, sizeof(int), hipMemcpyDeviceToHost, s->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    if (first != big) {
        const int i = first % n[0], j = (first / n[0]) % n[1], k = first / (n[0] * n[1]);
        set_error("node %d is not at grid point (%d,%d,%d)", first, i, j, k);
        return SFH_EARG;
    }
    for (int a = 0; a < d; ++a) {
        SFH_TRY(s->gaxis[a].ensure(sizeof(double) * n[a]));
        SFH_CHECK_HIP(hipMemcpyAsync(s->gaxis[a].p, g[a], sizeof(double) * n[a], hipMemcpyHostToDevice, s->ctx->stream));
        s->gn[a] = n[a];
    }
    s->dia_state = 0;
    s->gcell_state = 0;
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SFH_OK;
}

int sfh_assemble_matrix(sfh_system* s, int physics, int nq1d, const double* p1, const double* p2, double s1,
                        double s2, double factor) {
    if (s && s->kind == 0) { set_error("this system was built from triplets and has no mesh"); return SFH_ESTATE; }
    if (!s) { set_error("null system"); return SFH_EARG; }
    SFH_TRY(assemble_matrix(s, physics, nq1d, p1, p2, s1, s2, factor));
    s->nq1d = nq1d;
    return SFH_OK;
}

int sfh_assemble_rhs(sfh_system* s, int nq1d, const double* fconst, const double* fnodal, double factor, int accumulate) {
    if (s && s->kind == 0) { set_error("this system was built from triplets and has no mesh"); return SFH_ESTATE; }
    if (!s) { set_error("null system"); return SFH_EARG; }
    return assemble_rhs(s, nq1d, fconst, fnodal, factor, accumulate, s->F.as<double>());
}

int sfh_lumped_mass(sfh_system* s, int nq1d, double* out) {
    if (s && s->kind == 0) { set_error("this system was built from triplets and has no mesh"); return SFH_ESTATE; }
    if (!s || !out) { set_error("null argument"); return SFH_EARG; }
    const double one[3] = {1.0, 1.0, 1.0};
    return assemble_rhs(s, nq1d, one, nullptr, 1.0, 0, out);
}

int sfh_apply_constraints(sfh_system* s, int nc, const int* dofs, const double* values, double factor, double scalar) {
    if (!s) { set_error("null system"); return SFH_EARG; }
    if (!s->assembled) { set_error("constraints before assembly"); return SFH_ESTATE; }
    if (nc < 0 || (nc > 0 && !dofs)) { set_error("bad constraint list"); return SFH_EARG; }
    SFH_TRY(apply_constraints(s, nc, dofs, values, factor, scalar));
    s->nc += nc;
    return SFH_OK;
}

int sfh_clear_matrix(sfh_system* s) {
    SFH_CHECK_HIP(hipMemsetAsync(s->vals.p, 0, sizeof(double) * (size_t)s->nnz, s->ctx->stream));
    SFH_CHECK_HIP(hipMemsetAsync(s->vals32.p, 0, sizeof(float) * (size_t)s->nnz, s->ctx->stream));
    SFH_CHECK_HIP(hipMemsetAsync(s->cmask.p, 0, s->ndof, s->ctx->stream));
    s->nc = 0;
    s->dirty = true;
    if (s->dia_state > 0) s->dia_state = 0;
    return SFH_OK;
}

int sfh_clear_rhs(sfh_system* s) {
    SFH_CHECK_HIP(hipMemsetAsync(s->F.p, 0, sizeof(double) * s->ndof, s->ctx->stream));
    return SFH_OK;
}

int sfh_matrix_values(sfh_system* s, double* vals) {
    SFH_CHECK_HIP(hipMemcpyAsync(vals, s->vals.p, sizeof(double) * (size_t)s->nnz, hipMemcpyDeviceToHost, s->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SFH_OK;
}

int sfh_rhs(sfh_system* s, double* F) {
    SFH_CHECK_HIP(hipMemcpyAsync(F, s->F.p, sizeof(double) * s->ndof, hipMemcpyDeviceToHost, s->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SFH_OK;
}

int sfh_system_device_ptrs(sfh_system* s, const int** rp, const int** ci, const double** va, const double** F) {
    if (rp) *rp = s->rowptr.as<int>();
    if (ci) *ci = s->colind.as<int>();
    if (va) *va = s->vals.as<double>();
    if (F) *F = s->F.as<double>();
    return SFH_OK;
}

int sfh_solve(sfh_system* s, int trans, int k, const double* B, double* X, const sfh_solve_opts* o, sfh_solve_stats* st) {
    if (!s) { set_error("null system"); return SFH_EARG; }
    return solve(s, trans, k, B, k, X, k, o, st);
}

int sfh_system_set_segment_grids(sfh_system* s, int nseg, const int* seg_rows, const int* nx, const int* ny) {
    if (!s || nseg <= 0 || !seg_rows || !nx || !ny) { set_error("bad argument"); return SFH_EARG; }
    if (s->dim != 2 || s->dofpn != 1) { set_error("segment grids need a scalar 2D system"); return SFH_EARG; }
    if (seg_rows[0] != 0 || seg_rows[nseg] != s->ndof) { set_error("segments must cover all rows"); return SFH_EARG; }
    std::vector<int> g(3 * (size_t)nseg);
    for (int q = 0; q < nseg; ++q) {
        if (nx[q] < 1 || ny[q] < 1 || (long long)nx[q] * ny[q] != seg_rows[q + 1] - seg_rows[q]) {
            set_error("grid %d (%d x %d) does not match its %d rows", q, nx[q], ny[q], seg_rows[q + 1] - seg_rows[q]);
            return SFH_EARG;
        }
        g[3 * q] = seg_rows[q];
        g[3 * q + 1] = nx[q];
        g[3 * q + 2] = ny[q];
    }
    s->sgrid = g;
    if (s->dia_state != 0) s->dia_state = 0;
    return SFH_OK;
}

int sfh_solve_batched(sfh_system* s, int nseg, const int* seg_rows, int trans, int k, const double* B, double* X,
                      const sfh_solve_opts* o, sfh_solve_stats* st) {
    if (!s || nseg <= 0 || !seg_rows) { set_error("bad argument"); return SFH_EARG; }
    return solve(s, trans, k, B, k, X, k, o, st, nseg, seg_rows);
}

int sfh_solve_rhs(sfh_system* s, double* X, const sfh_solve_opts* o, sfh_solve_stats* st) {
    if (!s) { set_error("null system"); return SFH_EARG; }
    return solve(s, 0, 1, s->F.as<double>(), 1, X, 1, o, st);
}

int sfh_kernel_sampling(sfh_ctx* c, int every, int max_samples, int min_rhs) {
    if (!c || every < 0 || max_samples < 0 || min_rhs < 0) { set_error("bad argument"); return SFH_EARG; }
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    const size_t want = every > 0 ? 4 * (size_t)max_samples : 0;
    while (c->kev.size() > want) {
        (void)hipEventDestroy(c->kev.back());
        c->kev.pop_back();
    }
    while (c->kev.size() < want) {
        hipEvent_t e;
        SFH_CHECK_HIP(hipEventCreate(&e));
        c->kev.push_back(e);
    }
    c->ksample = every;
    c->kmin_rhs = min_rhs;
    c->ksum[0] = c->ksum[1] = c->ksum[2] = 0.0;
    c->kcount = 0;
    c->asum = 0.0;
    c->acount = 0;
    return SFH_OK;
}

int sfh_assembly_times(sfh_ctx* c, double* ms, long long* n) {
    if (!c || !ms) { set_error("bad argument"); return SFH_EARG; }
    *ms = c->acount > 0 ? c->asum / (double)c->acount : 0.0;
    if (n) *n = c->acount;
    return SFH_OK;
}

int sfh_kernel_times(sfh_ctx* c, double* ms, long long* n) {
    if (!c || !ms) { set_error("bad argument"); return SFH_EARG; }
    const double k = c->kcount > 0 ? (double)c->kcount : 1.0;
    for (int j = 0; j < 3; ++j) ms[j] = c->ksum[j] / k;
    if (n) *n = c->kcount;
    return SFH_OK;
}

// ------------------------------------------------------------ statistics
int sfh_covforcing_apply(sfh_system* s, const double* cv, double sigma, double ell, int k, const double* W, double* V,
                         int mode) {
    if (s && s->kind == 0) { set_error("this system was built from triplets and has no mesh"); return SFH_ESTATE; }
    if (!s || !cv || !W || !V || k <= 0) { set_error("bad argument"); return SFH_EARG; }
    return covforcing_apply(s, cv, sigma, ell, k, W, V, mode);
}

int sfh_gram(sfh_ctx* c, int N, int m, int n, const double* X, int ldx, const double* V, int ldv, double* C) {
    if (!c || N <= 0 || m <= 0 || n <= 0) { set_error("bad argument"); return SFH_EARG; }
    DevBuf work;
    SFH_TRY(gram(c, N, m, n, X, ldx, V, ldv, C, work));
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SFH_OK;
}

int sfh_gram_segmented(sfh_ctx* c, int nseg, const int* seg_rows, int m, int n, const double* X, int ldx,
                       const double* V, int ldv, double* C) {
    if (!c || nseg <= 0 || !seg_rows || m <= 0 || n <= 0 || !X || !V || !C) { set_error("bad argument"); return SFH_EARG; }
    for (int q = 0; q < nseg; ++q)
        if (seg_rows[q + 1] < seg_rows[q] || seg_rows[q] < 0) { set_error("segment bounds must ascend"); return SFH_EARG; }
    return gram_segmented(c, nseg, seg_rows, m, n, X, ldx, V, ldv, C);
}

int sfh_covforcing_grid(sfh_ctx* c, int nx, int ny, int nz, const double* gx, const double* gy, const double* gz,
                        const double* cv, double sigma, double ell, int k, const double* W, double* V) {
    if (!c || nx <= 0 || ny <= 0 || nz <= 0 || !gx || (ny > 1 && !gy) || (nz > 1 && !gz) || !cv || k <= 0 || !W ||
        !V || !(ell > 0)) {
        set_error("bad argument");
        return SFH_EARG;
    }
    const int d = nz > 1 ? 3 : (ny > 1 ? 2 : 1);
    const int gn[3] = {nx, ny, nz};
    DevBuf axes, sc0, sc1, sc2;
    SFH_TRY(axes.ensure(sizeof(double) * ((size_t)nx + ny + nz)));
    double* a = axes.as<double>();
    SFH_CHECK_HIP(hipMemcpyAsync(a, gx, sizeof(double) * nx, hipMemcpyHostToDevice, c->stream));
    if (d > 1) SFH_CHECK_HIP(hipMemcpyAsync(a + nx, gy, sizeof(double) * ny, hipMemcpyHostToDevice, c->stream));
    if (d > 2) SFH_CHECK_HIP(hipMemcpyAsync(a + nx + ny, gz, sizeof(double) * nz, hipMemcpyHostToDevice, c->stream));
    const double* ax[3] = {a, a + nx, a + nx + ny};
    SFH_TRY(covforcing_grid(c, d, gn, ax, cv, sigma, ell, k, W, V, sc0, sc1, sc2));
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SFH_OK;
}

int sfh_covforcing_batched(sfh_ctx* c, int nseg, const int* seg_rows, const int* nx, const int* ny,
                           const double* axes, const double* cv, const double* sigma, const double* ell, int k,
                           const double* W, double* V) {
    if (!c || nseg <= 0 || !seg_rows || !nx || !ny || !axes || !cv || !sigma || !ell || k <= 0 || !W || !V) {
        set_error("bad argument");
        return SFH_EARG;
    }
    for (int q = 0; q < nseg; ++q) {
        if (nx[q] <= 0 || ny[q] <= 0 || !(ell[q] > 0) || seg_rows[q] < 0 ||
            (long long)nx[q] * ny[q] != (long long)seg_rows[q + 1] - seg_rows[q]) {
            set_error("problem %d: grid %d x %d does not match its rows", q, nx[q], ny[q]);
            return SFH_EARG;
        }
    }
    return covforcing_batched(c, nseg, seg_rows, nx, ny, axes, cv, sigma, ell, k, W, V);
}

int sfh_system_set_covvec(sfh_system* s, const double* covvec) {
    if (!s) { set_error("bad argument"); return SFH_EARG; }
    if (!covvec) {
        s->fg_user = false;
        return SFH_OK;
    }
    SFH_TRY(s->fg_user_cv.ensure((size_t)s->ndof * 8));
    SFH_CHECK_HIP(hipMemcpyAsync(s->fg_user_cv.p, covvec, (size_t)s->ndof * 8, hipMemcpyHostToDevice, s->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    s->fg_user = true;
    return SFH_OK;
}

int sfh_forward_gauss(sfh_system* s, int ns, int pn, const int* pcols, const double* pvals, double sigma_f,
                      double ell_f, double reg_eps, const sfh_solve_opts* opts, double* mean, double* Cu,
                      double* field, double* ubar, sfh_solve_stats* stats) {
    if (s && s->kind == 0) { set_error("this system was built from triplets and has no mesh"); return SFH_ESTATE; }
    if (!s || ns <= 0 || pn <= 0 || !pcols || !pvals) { set_error("bad argument"); return SFH_EARG; }
    if (s->dofpn != 1) { set_error("forward model needs scalar DOFs"); return SFH_EARG; }
    if (!s->assembled) { set_error("forward model before assembly"); return SFH_ESTATE; }
    for (int i = 0; i < ns * pn; ++i)
        if (pcols[i] < 0 || pcols[i] >= s->ndof) { set_error("projection column out of range"); return SFH_EARG; }
    hipStream_t st = s->ctx->stream;
    const int N = s->ndof;
    const size_t nk = (size_t)N * ns;
    hipEvent_t e0, e1;
    SFH_CHECK_HIP(hipEventCreate(&e0));
    SFH_CHECK_HIP(hipEventCreate(&e1));
    SFH_CHECK_HIP(hipEventRecord(e0, st));
    SFH_TRY(s->fg_pt.ensure(nk * 8));
    SFH_TRY(s->fg_rt.ensure(nk * 8));
    SFH_TRY(s->fg_v.ensure(nk * 8));
    SFH_TRY(s->fg_cv.ensure((size_t)N * 8));
    SFH_TRY(s->fg_cu.ensure((size_t)ns * ns * 8 + (size_t)ns * 8));
    SFH_TRY(s->fg_cols.ensure(sizeof(int) * ns * pn));
    SFH_TRY(s->fg_vals.ensure(sizeof(double) * ns * pn));
    double* Pt = s->fg_pt.as<double>();
    double* Rt = s->fg_rt.as<double>();
    double* V = s->fg_v.as<double>();
    double* cv = s->fg_cv.as<double>();
    double* dCu = s->fg_cu.as<double>();
    double* dmean = dCu + (size_t)ns * ns;
    // covVec = int phi, zero on constrained DOFs (CovarianceForcing.hpp:101-114),
    // or the caller's weights
    if (s->fg_user) {
        SFH_CHECK_HIP(hipMemcpyAsync(cv, s->fg_user_cv.p, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
    } else {
        const double one = 1.0;
        SFH_TRY(assemble_rhs(s, s->nq1d, &one, nullptr, 1.0, 0, cv));
        k_mask_zero<<<nb(N, 256), 256, 0, st>>>(N, s->cmask.as<unsigned char>(), cv);
    }
    // P^T (N x ns)
    SFH_CHECK_HIP(hipMemsetAsync(Pt, 0, nk * 8, st));
    SFH_CHECK_HIP(hipMemcpyAsync(s->fg_cols.p, pcols, sizeof(int) * ns * pn, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(s->fg_vals.p, pvals, sizeof(double) * ns * pn, hipMemcpyHostToDevice, st));
    k_scatter_pt<<<nb(ns, 64), 64, 0, st>>>(ns, pn, s->fg_cols.as<int>(), s->fg_vals.as<double>(), ns, Pt);
    SFH_LAUNCH_CHECK();
    sfh_solve_stats s1 = {}, s2 = {}, s3 = {};
    // ubar = A^-1 F: a one-column solve of latency-bound iterations, run on the
    // second stream from a helper thread while the multi-RHS solves run here
    struct Background {
        std::thread t;
        int rc = SFH_OK;
        std::string err;
        ~Background() { if (t.joinable()) t.join(); }
        int join() {
            if (t.joinable()) t.join();
            if (rc != SFH_OK) set_error("%s", err.c_str());
            return rc;
        }
    } bg;
    if (ubar) {
        SFH_TRY(prepare_operator(s));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        bg.t = std::thread([&bg, s, ubar, opts, &s3] {
            hipSetDevice(s->ctx->device);
            bg.rc = solve_on(s, s->ctx->stream2, s->work2, 0, 1, s->F.as<double>(), 1, ubar, 1, opts, &s3);
            if (bg.rc != SFH_OK) bg.err = sfh_last_error();
        });
    }
    // R^T = A^-T P^T (ForwardGaussRHS.hpp:105-112)
    SFH_TRY(solve(s, 1, ns, Pt, ns, Rt, ns, opts, &s1));
    // mean = R F (:124)
    SFH_TRY(gram(s->ctx, N, ns, 1, Rt, ns, s->F.as<double>(), 1, dmean, s->fg_gram));
    // V = C_f R^T, Cu = R C_f R^T (CovarianceForcing.ipp:37-109)
    SFH_TRY(covforcing_apply(s, cv, sigma_f, ell_f, ns, Rt, V, 0));
    SFH_TRY(gram(s->ctx, N, ns, ns, Rt, ns, V, ns, dCu, s->fg_gram));
    const double reg = sigma_f * reg_eps;
    if (field) {
        // C_u P^T = A^-1 C_f R^T + reg^2 P^T
        SFH_TRY(solve(s, 0, ns, V, ns, field, ns, opts, &s2));
        k_pt_axpy<<<nb(ns, 64), 64, 0, st>>>(ns, pn, s->fg_cols.as<int>(), reg * reg, Pt, ns, field);
        SFH_LAUNCH_CHECK();
    }
    SFH_TRY(bg.join());
    SFH_CHECK_HIP(hipEventRecord(e1, st));
    // C_u and the mean are adjacent on the device: one copy into page-locked memory
    SFH_TRY(s->fg_host.ensure(((size_t)ns * ns + ns) * 8));
    SFH_CHECK_HIP(hipMemcpyAsync(s->fg_host.p, dCu, ((size_t)ns * ns + ns) * 8, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    double* hCu = s->fg_host.as<double>();
    const double* hm = hCu + (size_t)ns * ns;
    float ms = 0.f;
    SFH_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // + (sigma_f reg_eps)^2 P P^T (ForwardGaussRHS.hpp:131-136), from the sparse
    // rows: the entries of every DOF that several sensor rows share, grouped by
    // sorting the (DOF, sensor) pairs
    {
        std::vector<std::pair<int, int>> e((size_t)ns * pn);
        for (int a = 0; a < ns * pn; ++a) e[a] = {pcols[a], a};
        std::sort(e.begin(), e.end());
        const double r2 = reg * reg;
        for (size_t i = 0; i < e.size();) {
            size_t j = i;
            while (j < e.size() && e[j].first == e[i].first) ++j;
            for (size_t u = i; u < j; ++u)
                for (size_t v = i; v < j; ++v)
                    hCu[(size_t)(e[u].second / pn) * ns + e[v].second / pn] +=
                        r2 * pvals[e[u].second] * pvals[e[v].second];
            i = j;
        }
    }
    if (Cu) std::memcpy(Cu, hCu, (size_t)ns * ns * 8);
    if (mean) std::memcpy(mean, hm, (size_t)ns * 8);
    if (stats) {
        stats->outer = std::max(s1.outer, std::max(s2.outer, s3.outer));
        stats->inner_iters = s1.inner_iters + s2.inner_iters + s3.inner_iters;
        stats->max_relres = std::max(s1.max_relres, std::max(s2.max_relres, s3.max_relres));
        stats->ms = ms;
        stats->converged = 1;
        stats->layout = s1.layout;
        stats->precond = s1.precond;
        stats->levels = s1.levels;
        stats->graph_iters = s1.graph_iters + s2.graph_iters + s3.graph_iters;
    }
    return SFH_OK;
}

int sfh_lik_create(sfh_ctx* c, int ns, int no, int dim, const double* xs, const double* Cu, const double* mean,
                   const double* Y, double eps, sfh_lik** out) {
    if (!c || !out || ns <= 0 || no <= 0 || dim <= 0 || !xs || !Cu || !mean || !Y) { set_error("bad argument"); return SFH_EARG; }
    *out = nullptr;
    sfh_lik* L = new sfh_lik;
    L->ctx = c;
    L->ns = ns;
    L->no = no;
    L->dim = dim;
    L->eps = eps;
    L->xs_h.assign(xs, xs + (size_t)ns * dim);
    L->Cu_h.assign(Cu, Cu + (size_t)ns * ns);
    L->mean_h.assign(mean, mean + ns);
    L->Y_h.assign(Y, Y + (size_t)no * ns);
    int rc = L->xs.ensure(8 * (size_t)ns * dim);
    if (!rc) rc = L->Cu.ensure(8 * (size_t)ns * ns);
    if (!rc) rc = L->mean.ensure(8 * (size_t)ns);
    if (!rc) rc = L->Y.ensure(8 * (size_t)no * ns);
    if (rc) { delete L; return rc; }
    hipStream_t st = c->stream;
    if (hipMemcpyAsync(L->xs.p, xs, 8 * (size_t)ns * dim, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(L->Cu.p, Cu, 8 * (size_t)ns * ns, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(L->mean.p, mean, 8 * (size_t)ns, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(L->Y.p, Y, 8 * (size_t)no * ns, hipMemcpyHostToDevice, st) || hipStreamSynchronize(st)) {
        set_error("upload failed");
        delete L;
        return SFH_EDEVICE;
    }
    *out = L;
    return SFH_OK;
}

int sfh_lik_destroy(sfh_lik* L) {
    if (!L) return SFH_OK;
    (void)hipStreamSynchronize(L->ctx->stream);
    delete L;
    return SFH_OK;
}

int sfh_lik_loglik_batch(sfh_lik* L, int M, const double* theta, double* out, int* status) {
    if (!L || M <= 0 || !theta || !out) { set_error("bad argument"); return SFH_EARG; }
    hipStream_t st = L->ctx->stream;
    SFH_TRY(L->theta.ensure(sizeof(double) * 3 * M));
    SFH_TRY(L->out.ensure(sizeof(double) * M));
    SFH_TRY(L->status.ensure(sizeof(int) * M));
    SFH_CHECK_HIP(hipMemcpyAsync(L->theta.p, theta, sizeof(double) * 3 * M, hipMemcpyHostToDevice, st));
    SFH_TRY(loglik_batch(L, M, L->theta.as<double>(), L->out.as<double>(), L->status.as<int>()));
    std::vector<int> sth(M);
    SFH_CHECK_HIP(hipMemcpyAsync(out, L->out.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(sth.data(), L->status.p, sizeof(int) * M, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (status) std::memcpy(status, sth.data(), sizeof(int) * M);
    return SFH_OK;
}

int sfh_lik_mismatch(sfh_lik* L, double sd, double ld, double* Ks) {
    if (!L || !Ks) { set_error("bad argument"); return SFH_EARG; }
    const int ns = L->ns, d = L->dim;
    for (int k = 0; k < ns; ++k)
        for (int l = 0; l < ns; ++l) {
            double sq = 0.0;
            for (int a = 0; a < d; ++a) {
                const double dd = L->xs_h[(size_t)k * d + a] - L->xs_h[(size_t)l * d + a];
                sq += dd * dd;
            }
            Ks[(size_t)k * ns + l] = sd * sd * std::exp(-sq / (2. * ld * ld)) + (k == l ? L->eps * L->eps : 0.0);
        }
    return SFH_OK;
}

static int post_solve(sfh_lik* L, double rho, double sd, double ld, DevBuf& S, DevBuf& B) {
    const int ns = L->ns;
    hipStream_t st = L->ctx->stream;
    if (!(rho != 0.0)) { set_error("rho must be non-zero"); return SFH_EARG; }
    SFH_TRY(S.ensure(8 * (size_t)ns * ns));
    SFH_TRY(B.ensure(8 * (size_t)ns * (ns + 1)));
    k_post_build<<<nb((long long)ns * ns, 256), 256, 0, st>>>(ns, L->no, L->dim, L->xs.as<double>(), L->Cu.as<double>(),
                                                             L->mean.as<double>(), L->Y.as<double>(), L->eps, rho, sd,
                                                             ld, S.as<double>(), B.as<double>());
    SFH_LAUNCH_CHECK();
    DevBuf dst;
    SFH_TRY(dst.ensure(sizeof(int)));
    SFH_TRY(dense_spd_solve(L->ctx, ns, S.as<double>(), ns + 1, B.as<double>(), dst.as<int>()));
    int h = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&h, dst.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (h != SFH_OK) { set_error("C_u + K_s/(rho^2 n_o) not SPD"); return h; }
    return SFH_OK;
}

int sfh_lik_post_proj(sfh_lik* L, double rho, double sd, double ld, double* mean, double* cov) {
    if (!L || !mean || !cov) { set_error("bad argument"); return SFH_EARG; }
    const int ns = L->ns;
    hipStream_t st = L->ctx->stream;
    DevBuf S, B, T, C, m;
    SFH_TRY(post_solve(L, rho, sd, ld, S, B));
    SFH_TRY(T.ensure(8 * (size_t)ns * (ns + 1)));
    SFH_TRY(C.ensure(8 * (size_t)ns * ns));
    SFH_TRY(m.ensure(8 * (size_t)ns));
    // T = Cu [S^-1 Cu | S^-1 (ybar/rho - mean)]
    SFH_TRY(dgemm_shared_a(L->ctx, ns, ns + 1, ns, L->Cu.as<double>(), ns, B.as<double>(), ns + 1, 0, T.as<double>(),
                           ns + 1, 0, 1));
    k_post_finish<<<nb((long long)ns * ns, 256), 256, 0, st>>>(ns, L->Cu.as<double>(), L->mean.as<double>(),
                                                              T.as<double>(), C.as<double>(), m.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemcpyAsync(cov, C.p, 8 * (size_t)ns * ns, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(mean, m.p, 8 * (size_t)ns, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int sfh_lik_post_field(sfh_lik* L, double rho, double sd, double ld, int N, const double* field, const double* ubar,
                       double* mean) {
    if (!L || N <= 0 || !field || !ubar || !mean) { set_error("bad argument"); return SFH_EARG; }
    const int ns = L->ns;
    hipStream_t st = L->ctx->stream;
    DevBuf S, B;
    SFH_TRY(post_solve(L, rho, sd, ld, S, B));
    // v = column ns of B; mean = ubar + field v  (Woodbury mean, LikelihoodKOH.ipp:491-503)
    SFH_TRY(dgemm_shared_a(L->ctx, N, 1, ns, field, ns, B.as<double>() + ns, ns + 1, 0, mean, 1, 0, 1));
    k_axpy<<<nb(N, 256), 256, 0, st>>>(N, 1.0, ubar, mean);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int sfh_lik_post_field_cov(sfh_lik* L, double rho, double sd, double ld, int N, const double* field,
                           const double* prior, int full, double* out) {
    if (!L || N <= 0 || !field || !prior || !out) { set_error("bad argument"); return SFH_EARG; }
    if (full && (long long)N * N > (1LL << 31)) { set_error("full field covariance too large (N = %d)", N); return SFH_EARG; }
    const int ns = L->ns;
    hipStream_t st = L->ctx->stream;
    if (!(rho != 0.0)) { set_error("rho must be non-zero"); return SFH_EARG; }
    // S = C_u + K_s / (rho^2 n_o) and its inverse from the Cholesky factor (LikelihoodKOH.ipp:466-474)
    DevBuf S, B, Si, H, dst;
    SFH_TRY(S.ensure(8 * (size_t)ns * ns));
    SFH_TRY(B.ensure(8 * (size_t)ns * (ns + 1)));
    SFH_TRY(Si.ensure(8 * (size_t)ns * ns));
    SFH_TRY(H.ensure(8 * (size_t)N * ns));
    SFH_TRY(dst.ensure(sizeof(int)));
    k_post_build<<<nb((long long)ns * ns, 256), 256, 0, st>>>(ns, L->no, L->dim, L->xs.as<double>(), L->Cu.as<double>(),
                                                             L->mean.as<double>(), L->Y.as<double>(), L->eps, rho, sd,
                                                             ld, S.as<double>(), B.as<double>());
    k_eye<<<nb((long long)ns * ns, 256), 256, 0, st>>>(ns, Si.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(dense_spd_solve(L->ctx, ns, S.as<double>(), ns, Si.as<double>(), dst.as<int>()));
    int h = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&h, dst.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (h != SFH_OK) { set_error("C_u + K_s/(rho^2 n_o) not SPD"); return h; }
    // H = (C_u P^T) S^-1 (N x ns); D C_u = C_u - H (P C_u) with P C_u = (C_u P^T)^T (:484-488)
    SFH_TRY(dgemm_shared_a(L->ctx, N, ns, ns, field, ns, Si.as<double>(), ns, 0, H.as<double>(), ns, 0, 1));
    if (!full) {
        k_post_var<<<nb(N, 4), 256, 0, st>>>(N, ns, field, H.as<double>(), prior, out);
        SFH_LAUNCH_CHECK();
    } else {
        DevBuf Ft, T;
        SFH_TRY(Ft.ensure(8 * (size_t)N * ns));
        SFH_TRY(T.ensure(8 * (size_t)N * N));
        k_transpose<<<dim3((ns + 31) / 32, (N + 31) / 32), 256, 0, st>>>(N, ns, field, Ft.as<double>());
        SFH_LAUNCH_CHECK();
        SFH_TRY(dgemm_shared_a(L->ctx, N, N, ns, H.as<double>(), ns, Ft.as<double>(), N, 0, T.as<double>(), N, 0, 1));
        k_sub<<<nb((long long)N * N, 256), 256, 0, st>>>((long long)N * N, prior, T.as<double>(), out);
        SFH_LAUNCH_CHECK();
    }
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int sfh_lik_logdet_batch(sfh_lik* L, int M, const double* theta, double* logdet, int* status) {
    if (!L || M <= 0 || !theta || !logdet) { set_error("bad argument"); return SFH_EARG; }
    hipStream_t st = L->ctx->stream;
    SFH_TRY(L->theta.ensure(sizeof(double) * 3 * M));
    SFH_TRY(L->out.ensure(sizeof(double) * M));
    SFH_TRY(L->status.ensure(sizeof(int) * M));
    DevBuf ld;
    SFH_TRY(ld.ensure(sizeof(double) * M));
    SFH_CHECK_HIP(hipMemcpyAsync(L->theta.p, theta, sizeof(double) * 3 * M, hipMemcpyHostToDevice, st));
    SFH_TRY(loglik_batch(L, M, L->theta.as<double>(), L->out.as<double>(), L->status.as<int>(), ld.as<double>()));
    SFH_CHECK_HIP(hipMemcpyAsync(logdet, ld.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    if (status) SFH_CHECK_HIP(hipMemcpyAsync(status, L->status.p, sizeof(int) * M, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int sfh_mcmc(sfh_lik* L, int chains, int steps, const double* theta0, const double* propsd, uint64_t seed,
             int prior_mode, const double* prior_mu, const double* prior_var, double* samples, double* acc) {
    if (!L || chains <= 0 || steps <= 0 || !theta0 || !propsd) { set_error("bad argument"); return SFH_EARG; }
    if (prior_mode < 0 || prior_mode > 2) { set_error("prior_mode must be 0, 1 or 2"); return SFH_EARG; }
    if (prior_mode && (!prior_mu || !prior_var)) { set_error("prior parameters missing"); return SFH_EARG; }
    return mcmc_run(L, chains, steps, theta0, propsd, seed, prior_mode, prior_mu, prior_var, samples, acc);
}

int sfh_lik_typical_readings(sfh_lik* L, double rho, double sd, double ld, int nr, uint64_t seed, double* Y) {
    if (!L || nr <= 0 || !Y) { set_error("bad argument"); return SFH_EARG; }
    return typical_readings(L, rho, sd, ld, nr, seed, Y);
}

int sfh_spmv(sfh_system* s, int k, const double* X, double* Y) {
    if (!s || k <= 0 || !X || !Y) { set_error("bad argument"); return SFH_EARG; }
    SFH_TRY(spmv(s, k, X, Y));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SFH_OK;
}

int sfh_matern_tau(double sigma, double kappa, double nu, int dim, double* tau) {
    if (!(sigma > 0.0) || !(kappa > 0.0) || !(nu > 0.0) || dim < 1 || dim > 3 || !tau) {
        set_error("Matern hyperparameters must be positive");
        return SFH_EARG;
    }
    const double t2 = std::tgamma(nu) / ((sigma * sigma) * std::tgamma(nu + dim / 2.) *
                                          std::pow(4. * M_PI, dim / 2.) * std::pow(kappa, 2. * nu));
    *tau = std::sqrt(t2);
    return SFH_OK;
}

int sfh_precision_apply(sfh_system* s, const double* mass, double sigma, double kappa, double nu, int k,
                        const double* X, double* Y) {
    if (!s || !mass || k <= 0 || !X || !Y || s->dofpn != 1) { set_error("bad argument"); return SFH_EARG; }
    double tau = 0.0;
    SFH_TRY(sfh_matern_tau(sigma, kappa, nu, s->dim, &tau));
    const double beta = (nu + s->dim / 2.) / 2.;
    const double alpha = std::max(1., std::floor(beta));
    if (std::fabs(beta - alpha) >= 1e-12) {
        set_error("fractional SPDE exponent (beta = %g) needs the rational approximation", beta);
        return SFH_EARG;
    }
    return precision_apply(s, mass, tau * tau * std::pow(kappa, 4. * beta), (int)alpha, k, X, Y);
}

int sfh_frac_apply(sfh_system* s, const double* mass, int degree, const double* roots, int transpose, int k,
                   const double* X, double* Y) {
    if (!s || !mass || degree < 0 || degree > 32 || !roots || k <= 0 || !X || !Y || s->dofpn != 1) {
        set_error("bad argument");
        return SFH_EARG;
    }
    return frac_apply(s, mass, degree, roots, transpose ? 1 : 0, k, X, Y);
}

int sfh_precision_apply_rational(sfh_system* s, const double* mass, double sigma, double kappa, double nu,
                                 int degree, const double* bot, int k, const double* X, double* Y) {
    if (!s || !mass || degree < 0 || degree > 32 || !bot || k <= 0 || !X || !Y || s->dofpn != 1) {
        set_error("bad argument");
        return SFH_EARG;
    }
    double tau = 0.0;
    SFH_TRY(sfh_matern_tau(sigma, kappa, nu, s->dim, &tau));
    const double beta = (nu + s->dim / 2.) / 2.;
    const double alpha = std::max(1., std::floor(beta));
    const size_t total = (size_t)s->ndof * k;
    DevBuf a, b;
    SFH_TRY(a.ensure(sizeof(double) * total));
    SFH_TRY(b.ensure(sizeof(double) * total));
    SFH_TRY(frac_apply(s, mass, degree, bot, 0, k, X, a.as<double>()));              // P_l X
    SFH_TRY(precision_apply(s, mass, tau * tau * std::pow(kappa, 4. * beta), (int)alpha, k, a.as<double>(),
                            b.as<double>()));                                         // Q P_l X
    return frac_apply(s, mass, degree, bot, 1, k, b.as<double>(), Y);                 // P_l^T Q P_l X
}

int sfh_gp(sfh_ctx* c, int n, int dim, const double* X, const double* y, double sigma, double ell, double eps, int np,
           const double* Xp, double* logl, double* mean, double* var) {
    if (!c || n <= 0 || dim <= 0 || !X || !y || !logl || (np > 0 && (!Xp || !mean || !var))) {
        set_error("bad argument");
        return SFH_EARG;
    }
    return gp_run(c, n, dim, X, y, sigma, ell, eps, np, Xp, logl, mean, var, nullptr);
}

int sfh_gp_noalloc(sfh_ctx* c, int n, int dim, const double* X, const double* y, double sigma, double ell, double eps,
                   int np, const double* Xp, double* logl, double* mean, double* var, int* status) {
    if (!c || n <= 0 || dim <= 0 || !X || !y || !logl || !status || np < 0 || (np > 0 && (!Xp || !mean || !var))) {
        set_error("bad argument");
        return SFH_EARG;
    }
    if (n > kGpFixedN || np > kGpFixedNP) {
        set_error("sfh_gp_noalloc takes at most %d observations and %d prediction points (got %d, %d)", kGpFixedN,
                  kGpFixedNP, n, np);
        return SFH_EARG;
    }
    if (!c->gp_scratch) { set_error("no GP scratch on device %d", c->device); return SFH_ESTATE; }
    return gp_run_fixed(c, n, dim, X, y, sigma, ell, eps, np, Xp, logl, mean, var, status);
}

int sfh_gp_predict_cov(sfh_ctx* c, int n, int dim, const double* X, const double* y, double sigma, double ell,
                       double eps, int np, const double* Xp, double* cov) {
    if (!c || n <= 0 || dim <= 0 || !X || !y || np <= 0 || !Xp || !cov) {
        set_error("bad argument");
        return SFH_EARG;
    }
    std::vector<double> m(np), v(np);
    double logl = 0.0;
    return gp_run(c, n, dim, X, y, sigma, ell, eps, np, Xp, &logl, m.data(), v.data(), cov);
}

}  // extern "C"
