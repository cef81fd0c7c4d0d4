// fields.hip -- device kernels around the pressure solve: divergence of the
// staggered velocity (RHS of MG_MODE and the continuity error), the PCG
// coefficient beta = alpha dt / rho, the MG_MODE projection, and the FAST_MODE
// source and projection of the constant-coefficient splitting (reference
// density rho0, extrapolated old pressure 2 psi^n - psi^(n-1)).
//
// Staggered layout: ux, uz hold nx*ny*nz faces (face i between cells i and i+1,
// face 0 == face nx by periodicity); uy holds nx*(ny+1)*nz faces, j = 0..ny,
// with the walls at j = 0 and j = ny.
#include <cstring>
#include <map>
#include <mutex>

#include "fields.hpp"

namespace tbfg {

// per-axis metrics: df[i-1] = df(i), dcn[i-1] = dc(i+1), i = 1..n
struct FieldGeo {
    int nx, ny, nz;
    const double* df[3];
    const double* dcn[3];
};

struct FieldGeoOwner {
    DevBuf buf;
    FieldGeo view{};
};

static std::mutex g_geo_mu;
static std::map<std::string, std::unique_ptr<FieldGeoOwner>> g_geo_cache;

static int get_geo(const tbfg_grid_desc* g, FieldGeo& out)
{
    TBFG_TRY(validate_grid(g));
    int dev = 0;
    TBFG_HIP(hipGetDevice(&dev));
    const std::string key = grid_key(g, dev);
    std::lock_guard<std::mutex> lk(g_geo_mu);
    auto it = g_geo_cache.find(key);
    if (it == g_geo_cache.end()) {
        HostLevel L = make_level(g, 0);
        std::vector<double> h;
        size_t off[3][2];
        for (int a = 0; a < 3; ++a) {
            off[a][0] = h.size();
            for (int i = 1; i <= L.n[a]; ++i) h.push_back(L.ax[a].df(i));
            off[a][1] = h.size();
            for (int i = 1; i <= L.n[a]; ++i) h.push_back(L.ax[a].dc(i + 1));
        }
        std::unique_ptr<FieldGeoOwner> o(new FieldGeoOwner());
        TBFG_TRY(upload(o->buf, h.data(), h.size() * 8));
        o->view.nx = L.n[0];
        o->view.ny = L.n[1];
        o->view.nz = L.n[2];
        for (int a = 0; a < 3; ++a) {
            o->view.df[a] = o->buf.as<double>() + off[a][0];
            o->view.dcn[a] = o->buf.as<double>() + off[a][1];
        }
        it = g_geo_cache.emplace(key, std::move(o)).first;
    }
    out = it->second->view;
    return TBFG_OK;
}

// (i, j, k) of a thread in a row-wise launch made as a 1-D grid of
// ceil(nx / blockDim.x) * ny * nz workgroups.  Workgroups go round-robin to the 8
// XCDs, each with its own L2, so consecutive ones are renumbered to fall on one
// XCD: each XCD then works through a contiguous range of rows and planes, and the
// y and z neighbours a workgroup reads were read into that XCD's L2 by its
// neighbours (a 3-D grid numbered rows first sends neighbouring rows to different
// XCDs, and each of them fetches the shared faces from HBM again).
__device__ __forceinline__ void row_cell(int nx, int ny, int& i, int& j, int& k)
{
    const unsigned nb = gridDim.x;
    unsigned id = blockIdx.x;
    if ((nb & 7u) == 0) id = (id & 7u) * (nb >> 3) + (id >> 3);
    const unsigned gx = ((unsigned)nx + blockDim.x - 1) / blockDim.x;
    const unsigned bx = id % gx, r = id / gx;
    i = (int)(bx * blockDim.x + threadIdx.x);
    j = (int)(r % (unsigned)ny);
    k = (int)(r / (unsigned)ny);
}

__device__ __forceinline__ double div_cell(const FieldGeo& G, const double* __restrict__ ux,
                                           const double* __restrict__ uy, const double* __restrict__ uz, int i, int j,
                                           int k)
{
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const size_t c = i + (size_t)nx * (j + (size_t)ny * k);
    const size_t w = (i > 0 ? i - 1 : nx - 1) + (size_t)nx * (j + (size_t)ny * k);
    const size_t b = i + (size_t)nx * (j + (size_t)ny * (k > 0 ? k - 1 : nz - 1));
    const size_t yt = i + (size_t)nx * ((j + 1) + (size_t)(ny + 1) * k);
    const size_t yb = i + (size_t)nx * (j + (size_t)(ny + 1) * k);
    const double dx = (ux[c] - ux[w]) / G.df[0][i];
    const double dy = (uy[yt] - uy[yb]) / G.df[1][j];
    const double dz = (uz[c] - uz[b]) / G.df[2][k];
    return dx + dy + dz;
}

// poissonEqn.f90:176-188 / auxiliaryRoutines.f90:66-79
__global__ void k_divergence(FieldGeo G, const double* __restrict__ ux, const double* __restrict__ uy,
                             const double* __restrict__ uz, double* __restrict__ div)
{
    int i, j, k;
    row_cell(G.nx, G.ny, i, j, k);
    if (i >= G.nx) return;
    div[i + (size_t)G.nx * (j + (size_t)G.ny * k)] = div_cell(G, ux, uy, uz, i, j, k);
}

// auxiliaryRoutines.f90:60-82: max |div u| (the reference prints it times dt)
__global__ void k_maxdiv(FieldGeo G, const double* __restrict__ ux, const double* __restrict__ uy,
                         const double* __restrict__ uz, unsigned long long* out)
{
    int i, j, k;
    row_cell(G.nx, G.ny, i, j, k);
    double v = 0.0;
    if (i < G.nx) v = fabs(div_cell(G, ux, uy, uz, i, j, k));
    if (!(v == v)) v = __longlong_as_double(0x7ff0000000000000LL);  // NaN -> +inf
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    __shared__ double wmax[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wmax[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) m = fmax(m, wmax[w]);
        // non-negative doubles order like their bit patterns
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// poissonEqn.f90:331-337
__global__ void k_beta(size_t n, double alpha_dt, const double* __restrict__ rho, double* __restrict__ beta)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) beta[t] = alpha_dt / rho[t];
}

// momentumEqn.f90:806-866 (addPressureGrad) on all periodic x/z faces and the
// interior y faces (setEqnBounds, momentumEqn.f90:1134-1190).
__global__ void k_project_mg(FieldGeo G, double alpha_dt, const double* __restrict__ rho,
                             const double* __restrict__ psi, double* __restrict__ ux, double* __restrict__ uy,
                             double* __restrict__ uz)
{
    int i, j, k;
    row_cell(G.nx, G.ny, i, j, k);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    if (i >= nx) return;
    const size_t c = i + (size_t)nx * (j + (size_t)ny * k);
    const size_t e = (i + 1 < nx ? i + 1 : 0) + (size_t)nx * (j + (size_t)ny * k);
    const size_t f = i + (size_t)nx * (j + (size_t)ny * (k + 1 < nz ? k + 1 : 0));
    const double ic = 1.0 / rho[c];
    double inv = 0.5 * (1.0 / rho[e] + ic);
    ux[c] = ux[c] - alpha_dt * inv * (psi[e] - psi[c]) / G.dcn[0][i];
    inv = 0.5 * (1.0 / rho[f] + ic);
    uz[c] = uz[c] - alpha_dt * inv * (psi[f] - psi[c]) / G.dcn[2][k];
    if (j + 1 < ny) {
        const size_t n = i + (size_t)nx * ((j + 1) + (size_t)ny * k);
        const size_t yf = i + (size_t)nx * ((j + 1) + (size_t)(ny + 1) * k);
        inv = 0.5 * (1.0 / rho[n] + ic);
        uy[yf] = uy[yf] - alpha_dt * inv * (psi[n] - psi[c]) / G.dcn[1][j];
    }
}

// old-pressure level, absent on the first steps (then zero)
__device__ __forceinline__ double old_level(const double* __restrict__ a, size_t i) { return a ? a[i] : 0.0; }

// extrapolated old-pressure difference across a face: (2 (psi1_b - psi1_a) - (psi2_b - psi2_a)) / dc
__device__ __forceinline__ double old_grad(const double* __restrict__ psi1, const double* __restrict__ psi2, size_t a,
                                           size_t b, double dc)
{
    return (2.0 * (old_level(psi1, b) - old_level(psi1, a)) - (old_level(psi2, b) - old_level(psi2, a))) / dc;
}

// Old surface-tension levels st^n (1) and st^(n-1) (2), three staggered
// components laid out like the velocity; a NULL component is zero.
struct StLevels {
    const double* x1;
    const double* y1;
    const double* z1;
    const double* x2;
    const double* y2;
    const double* z2;
};

// the face value G_f = old_grad - 2 st^n_f + st^(n-1)_f (poissonEqn.f90:255-301,
// momentumEqn.f90:1104-1121), in the reference's order of operations
__device__ __forceinline__ double with_st(double grad, const double* __restrict__ s1, const double* __restrict__ s2,
                                          size_t f)
{
    return grad - 2.0 * old_level(s1, f) + old_level(s2, f);
}

// FAST computeSource + computeOldPressDiv with two old levels (poissonEqn.f90:193-305):
// s = rho0/(alpha dt) div u* + div[(1 - rho0 (1/rho)_f) G_f], G_f the face gradient of
// 2 psi^n - psi^(n-1) minus 2 st^n_f - st^(n-1)_f; (1/rho)_f the arithmetic face mean.
// The face after a cell reads the st values of the cell's own face index, the face
// before it those of the previous face (in y the wall faces 0 and ny included).
__global__ void k_source_fast(FieldGeo G, double r_adt, double rho0, const double* __restrict__ rho,
                              const double* __restrict__ ux, const double* __restrict__ uy,
                              const double* __restrict__ uz, const double* __restrict__ psi1,
                              const double* __restrict__ psi2, StLevels st, double* __restrict__ s)
{
    int i, j, k;
    row_cell(G.nx, G.ny, i, j, k);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    if (i >= nx) return;
    const size_t plane = (size_t)nx * ny;
    const size_t c = i + (size_t)nx * j + plane * k;
    const size_t e = (i + 1 < nx ? i + 1 : 0) + (size_t)nx * j + plane * k;
    const size_t w = (i > 0 ? i - 1 : nx - 1) + (size_t)nx * j + plane * k;
    const size_t n = i + (size_t)nx * (j + 1 < ny ? j + 1 : j) + plane * k;  // wall ghost = the cell
    const size_t so = i + (size_t)nx * (j > 0 ? j - 1 : j) + plane * k;
    const size_t f = i + (size_t)nx * j + plane * (k + 1 < nz ? k + 1 : 0);
    const size_t b = i + (size_t)nx * j + plane * (k > 0 ? k - 1 : nz - 1);
    const size_t yt = i + (size_t)nx * ((j + 1) + (size_t)(ny + 1) * k);  // y face after the cell
    const size_t yb = yt - nx;                                             // y face before it
    const double dxf = G.df[0][0], dzf = G.df[2][0], dyf = G.df[1][j];
    const double dxc = G.dcn[0][0], dzc = G.dcn[2][0];
    const double dycp = G.dcn[1][j], dycm = G.dcn[1][j > 0 ? j - 1 : 0];
    const double sv = div_cell(G, ux, uy, uz, i, j, k) * r_adt * rho0;
    const double ic = 1.0 / rho[c];
    const double apx = 1.0 - rho0 * 0.5 * (1.0 / rho[e] + ic);
    const double amx = 1.0 - rho0 * 0.5 * (1.0 / rho[w] + ic);
    const double apy = 1.0 - rho0 * 0.5 * (1.0 / rho[n] + ic);
    const double amy = 1.0 - rho0 * 0.5 * (1.0 / rho[so] + ic);
    const double apz = 1.0 - rho0 * 0.5 * (1.0 / rho[f] + ic);
    const double amz = 1.0 - rho0 * 0.5 * (1.0 / rho[b] + ic);
    const double gxp = with_st(old_grad(psi1, psi2, c, e, dxc), st.x1, st.x2, c);
    const double gxm = with_st(old_grad(psi1, psi2, w, c, dxc), st.x1, st.x2, w);
    const double gyp = with_st(old_grad(psi1, psi2, c, n, dycp), st.y1, st.y2, yt);
    const double gym = with_st(old_grad(psi1, psi2, so, c, dycm), st.y1, st.y2, yb);
    const double gzp = with_st(old_grad(psi1, psi2, c, f, dzc), st.z1, st.z2, c);
    const double gzm = with_st(old_grad(psi1, psi2, b, c, dzc), st.z1, st.z2, b);
    const double dx = (apx * gxp - amx * gxm) / dxf;
    const double dy = (apy * gyp - amy * gym) / dyf;
    const double dz = (apz * gzp - amz * gzm) / dzf;
    s[c] = sv + dx + dy + dz;
}

// FAST makeVelocityDivFree + computeOldPressGrad (momentumEqn.f90:999-1127):
// u_f -= alpha dt [(1/rho0) d psi + ((1/rho)_f - 1/rho0) G_f] on the periodic x/z
// faces and the interior y faces.
__global__ void k_project_fast(FieldGeo G, double alpha_dt, double rho0, const double* __restrict__ rho,
                               const double* __restrict__ psi, const double* __restrict__ psi1,
                               const double* __restrict__ psi2, StLevels st, double* __restrict__ ux,
                               double* __restrict__ uy, double* __restrict__ uz)
{
    int i, j, k;
    row_cell(G.nx, G.ny, i, j, k);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    if (i >= nx) return;
    const size_t plane = (size_t)nx * ny;
    const size_t c = i + (size_t)nx * j + plane * k;
    const size_t e = (i + 1 < nx ? i + 1 : 0) + (size_t)nx * j + plane * k;
    const size_t f = i + (size_t)nx * j + plane * (k + 1 < nz ? k + 1 : 0);
    const double dxc = G.dcn[0][0], dzc = G.dcn[2][0];
    const double inv0 = 1.0 / rho0;
    const double ic = 1.0 / rho[c];
    double inv = 0.5 * (1.0 / rho[e] + ic);
    ux[c] = ux[c] - alpha_dt * (inv0 * ((psi[e] - psi[c]) / dxc) +
                                (inv - inv0) * with_st(old_grad(psi1, psi2, c, e, dxc), st.x1, st.x2, c));
    inv = 0.5 * (1.0 / rho[f] + ic);
    uz[c] = uz[c] - alpha_dt * (inv0 * ((psi[f] - psi[c]) / dzc) +
                                (inv - inv0) * with_st(old_grad(psi1, psi2, c, f, dzc), st.z1, st.z2, c));
    if (j + 1 < ny) {
        const size_t n = c + nx;
        const size_t yf = i + (size_t)nx * ((j + 1) + (size_t)(ny + 1) * k);
        const double dyc = G.dcn[1][j];
        inv = 0.5 * (1.0 / rho[n] + ic);
        uy[yf] = uy[yf] - alpha_dt * (inv0 * ((psi[n] - psi[c]) / dyc) +
                                      (inv - inv0) * with_st(old_grad(psi1, psi2, c, n, dyc), st.y1, st.y2, yf));
    }
}

// updatePressure (poissonEqn.f90:374-415) + clean_st_prev_bc (:419-459) for the
// channel, one thread per face row entry: p = psi; psi2 <- psi1 <- psi
// (storeOldField); st2 <- st1 <- st (storeOldFieldV, st NULL = zero); the old st
// levels zeroed on the y wall faces (j = 0, ny).  Block row j = ny carries only
// the top wall face of the y component.  Each entry is read and written by its
// own thread, so the rotation is in place.
struct StRotate {
    const double* cur[3];
    double* l1[3];
    double* l2[3];
};

__global__ void k_update_pressure(int nx, int ny, int nz, const double* __restrict__ psi, double* __restrict__ p,
                                  double* __restrict__ psi1, double* __restrict__ psi2, StRotate st, int with_st)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i >= nx) return;
    if (j < ny) {
        const size_t c = i + (size_t)nx * (j + (size_t)ny * k);
        const double v = psi[c];
        if (p) p[c] = v;
        psi2[c] = psi1[c];
        psi1[c] = v;
        if (with_st) {
            for (int a = 0; a < 3; a += 2) {
                st.l2[a][c] = st.l1[a][c];
                st.l1[a][c] = st.cur[a] ? st.cur[a][c] : 0.0;
            }
        }
    }
    if (with_st) {
        const size_t yf = i + (size_t)nx * (j + (size_t)(ny + 1) * k);
        if (j == 0 || j == ny) {
            st.l2[1][yf] = 0.0;
            st.l1[1][yf] = 0.0;
        } else {
            st.l2[1][yf] = st.l1[1][yf];
            st.l1[1][yf] = st.cur[1] ? st.cur[1][yf] : 0.0;
        }
    }
}

static dim3 row_block(int nx) { return dim3(nx >= 256 ? 256 : (unsigned)next_pow2(nx < 64 ? 64 : nx)); }

int divergence_dev(const tbfg_grid_desc* g, const double* ux, const double* uy, const double* uz, double* div,
                   void* stream)
{
    if (!ux || !uy || !uz || !div) {
        set_error("tbfg_divergence_dev: NULL argument");
        return TBFG_EINVAL;
    }
    FieldGeo G;
    TBFG_TRY(get_geo(g, G));
    dim3 blk = row_block(G.nx);
    hipLaunchKernelGGL(k_divergence, dim3(ceil_div(G.nx, blk.x) * G.ny * G.nz), blk, 0, (hipStream_t)stream, G, ux, uy,
                       uz, div);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

int max_divergence_dev(const tbfg_grid_desc* g, const double* ux, const double* uy, const double* uz, double* out,
                       void* stream)
{
    if (!ux || !uy || !uz || !out) {
        set_error("tbfg_max_divergence_dev: NULL argument");
        return TBFG_EINVAL;
    }
    FieldGeo G;
    TBFG_TRY(get_geo(g, G));
    DevBuf res;
    TBFG_TRY(res.alloc(sizeof(unsigned long long)));
    hipStream_t s = (hipStream_t)stream;
    TBFG_HIP(hipMemsetAsync(res.p, 0, sizeof(unsigned long long), s));
    dim3 blk = row_block(G.nx);
    hipLaunchKernelGGL(k_maxdiv, dim3(ceil_div(G.nx, blk.x) * G.ny * G.nz), blk, 0, s, G, ux, uy, uz,
                       res.as<unsigned long long>());
    TBFG_HIP(hipGetLastError());
    unsigned long long bits = 0;
    TBFG_HIP(hipMemcpyAsync(&bits, res.p, sizeof bits, hipMemcpyDeviceToHost, s));
    TBFG_HIP(hipStreamSynchronize(s));
    double v;
    static_assert(sizeof v == sizeof bits, "64-bit double");
    std::memcpy(&v, &bits, sizeof v);
    *out = v;
    return TBFG_OK;
}

// ----------------------------------------------- continuity error without allocation
//
// The same kernel as max_divergence_dev with its scratch -- the max word and the
// per-axis metrics -- in one file-scope device array instead of a hipMalloc per call
// and the geometry cache: [max word, padding to 256 B | df, dc(i+1) of x | of y | of z].
// Each device holds its own copy; its address comes from hipGetSymbolAddress on the
// first call on that device (the array's name is not a device address in host code).
constexpr int kNoAllocMaxAxis = 4096;   // cells per axis
constexpr size_t kNoAllocHead = 32;     // doubles before the metrics
constexpr size_t kNoAllocDoubles = kNoAllocHead + 2 * 3 * (size_t)kNoAllocMaxAxis;  // 196,864 bytes
static __device__ __align__(256) double scratch[kNoAllocDoubles];

struct NoAllocScratch {
    std::mutex mu;           // held for a whole call: the call ends with its stream drained
    double* ptr = nullptr;   // this device's copy of `scratch`
};

static std::mutex g_noalloc_mu;
static std::map<int, std::unique_ptr<NoAllocScratch>> g_noalloc;  // by device

static int noalloc_scratch(int dev, NoAllocScratch** out)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    std::unique_ptr<NoAllocScratch>& slot = g_noalloc[dev];
    if (!slot) {
        void* p = nullptr;
        TBFG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        std::unique_ptr<NoAllocScratch> d(new NoAllocScratch());
        d->ptr = static_cast<double*>(p);
        slot = std::move(d);
    }
    *out = slot.get();
    return TBFG_OK;
}

int max_divergence_noalloc_dev(const tbfg_grid_desc* g, const double* ux, const double* uy, const double* uz,
                               double* out, void* stream)
{
    if (!ux || !uy || !uz || !out) {
        set_error("tbfg_max_divergence_noalloc_dev: NULL argument");
        return TBFG_EINVAL;
    }
    TBFG_TRY(validate_grid(g));
    for (int a = 0; a < 3; ++a)
        if (g->n[a] > kNoAllocMaxAxis) {
            set_error("tbfg_max_divergence_noalloc_dev: axis %d has %d cells, at most %d fit its scratch", a,
                      g->n[a], kNoAllocMaxAxis);
            return TBFG_EINVAL;
        }
    int dev = 0;
    TBFG_HIP(hipGetDevice(&dev));
    NoAllocScratch* d = nullptr;
    TBFG_TRY(noalloc_scratch(dev, &d));
    std::lock_guard<std::mutex> lk(d->mu);

    // host image of the used part of the scratch: a zero max word, then the metrics
    // (the values get_geo uploads, so both entry points compute the same result)
    const HostLevel L = make_level(g, 0);
    std::vector<double> h(kNoAllocHead, 0.0);
    FieldGeo G;
    G.nx = L.n[0];
    G.ny = L.n[1];
    G.nz = L.n[2];
    size_t off[3][2];
    for (int a = 0; a < 3; ++a) {
        off[a][0] = h.size();
        for (int i = 1; i <= L.n[a]; ++i) h.push_back(L.ax[a].df(i));
        off[a][1] = h.size();
        for (int i = 1; i <= L.n[a]; ++i) h.push_back(L.ax[a].dc(i + 1));
    }
    for (int a = 0; a < 3; ++a) {
        G.df[a] = d->ptr + off[a][0];
        G.dcn[a] = d->ptr + off[a][1];
    }
    hipStream_t s = (hipStream_t)stream;
    unsigned long long bits = 0;
    hipError_t e = hipMemcpyAsync(d->ptr, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const dim3 blk = row_block(G.nx);
        hipLaunchKernelGGL(k_maxdiv, dim3(ceil_div(G.nx, blk.x) * G.ny * G.nz), blk, 0, s, G, ux, uy, uz,
                           reinterpret_cast<unsigned long long*>(d->ptr));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bits, d->ptr, sizeof bits, hipMemcpyDeviceToHost, s);
    // drained on every path: `h` and `bits` live on this frame, and the next call reuses the scratch
    const hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) {
        set_error("tbfg_max_divergence_noalloc_dev: %s", hipGetErrorString(e));
        return TBFG_EDEVICE;
    }
    double v;
    std::memcpy(&v, &bits, sizeof v);
    *out = v;
    return TBFG_OK;
}

int beta_dev(const tbfg_grid_desc* g, double alpha_dt, const double* rho, double* beta, void* stream)
{
    TBFG_TRY(validate_grid(g));
    if (!rho || !beta) {
        set_error("tbfg_beta_dev: NULL argument");
        return TBFG_EINVAL;
    }
    const size_t n = (size_t)g->n[0] * g->n[1] * g->n[2];
    hipLaunchKernelGGL(k_beta, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, n, alpha_dt, rho, beta);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

// a NULL level, or a NULL component of one, is zero
static StLevels st_levels(const double* const st1[3], const double* const st2[3])
{
    return StLevels{st1 ? st1[0] : nullptr, st1 ? st1[1] : nullptr, st1 ? st1[2] : nullptr,
                    st2 ? st2[0] : nullptr, st2 ? st2[1] : nullptr, st2 ? st2[2] : nullptr};
}

int source_fast_dev(const tbfg_grid_desc* g, double alpha_dt, double rho0, const double* rho, const double* ux,
                    const double* uy, const double* uz, const double* psi1, const double* psi2,
                    const double* const st1[3], const double* const st2[3], double* s, void* stream)
{
    if (!rho || !ux || !uy || !uz || !s) {
        set_error("tbfg_source_fast_dev: NULL argument");
        return TBFG_EINVAL;
    }
    if (!(alpha_dt > 0.0) || !(rho0 > 0.0)) {
        set_error("tbfg_source_fast_dev: alpha_dt = %g, rho0 = %g (both must be > 0)", alpha_dt, rho0);
        return TBFG_EINVAL;
    }
    FieldGeo G;
    TBFG_TRY(get_geo(g, G));
    if (g->stretch[0] || g->stretch[2]) {
        set_error("tbfg_source_fast_dev: x and z must be uniform (FAST_MODE)");
        return TBFG_EINVAL;
    }
    const StLevels st = st_levels(st1, st2);
    dim3 blk = row_block(G.nx);
    hipLaunchKernelGGL(k_source_fast, dim3(ceil_div(G.nx, blk.x) * G.ny * G.nz), blk, 0, (hipStream_t)stream, G,
                       1.0 / alpha_dt, rho0, rho, ux, uy, uz, psi1, psi2, st, s);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

int project_fast_dev(const tbfg_grid_desc* g, double alpha_dt, double rho0, const double* rho, const double* psi,
                     const double* psi1, const double* psi2, const double* const st1[3],
                     const double* const st2[3], double* ux, double* uy, double* uz, void* stream)
{
    if (!rho || !psi || !ux || !uy || !uz) {
        set_error("tbfg_project_fast_dev: NULL argument");
        return TBFG_EINVAL;
    }
    if (!(rho0 > 0.0)) {
        set_error("tbfg_project_fast_dev: rho0 = %g (must be > 0)", rho0);
        return TBFG_EINVAL;
    }
    FieldGeo G;
    TBFG_TRY(get_geo(g, G));
    if (g->stretch[0] || g->stretch[2]) {
        set_error("tbfg_project_fast_dev: x and z must be uniform (FAST_MODE)");
        return TBFG_EINVAL;
    }
    const StLevels st = st_levels(st1, st2);
    dim3 blk = row_block(G.nx);
    hipLaunchKernelGGL(k_project_fast, dim3(ceil_div(G.nx, blk.x) * G.ny * G.nz), blk, 0, (hipStream_t)stream, G,
                       alpha_dt, rho0, rho, psi, psi1, psi2, st, ux, uy, uz);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

int update_pressure_dev(const tbfg_grid_desc* g, const double* psi, double* p, double* psi1, double* psi2,
                        const double* const st[3], double* const st1[3], double* const st2[3], void* stream)
{
    TBFG_TRY(validate_grid(g));
    if (!psi || !psi1 || !psi2) {
        set_error("tbfg_update_pressure_dev: psi, psi1 and psi2 must not be NULL");
        return TBFG_EINVAL;
    }
    if ((st1 == nullptr) != (st2 == nullptr)) {
        set_error("tbfg_update_pressure_dev: st1 and st2 must both be given or both be NULL");
        return TBFG_EINVAL;
    }
    StRotate r{};
    const int with_st = st1 != nullptr;
    if (with_st) {
        for (int a = 0; a < 3; ++a) {
            if (!st1[a] || !st2[a]) {
                set_error("tbfg_update_pressure_dev: st1[%d] or st2[%d] is NULL", a, a);
                return TBFG_EINVAL;
            }
            r.cur[a] = st ? st[a] : nullptr;
            r.l1[a] = st1[a];
            r.l2[a] = st2[a];
        }
    } else if (st) {
        set_error("tbfg_update_pressure_dev: st given without st1, st2 to store it in");
        return TBFG_EINVAL;
    }
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    dim3 blk = row_block(nx);
    hipLaunchKernelGGL(k_update_pressure, dim3(ceil_div(nx, blk.x), ny + 1, nz), blk, 0, (hipStream_t)stream, nx, ny,
                       nz, psi, p, psi1, psi2, r, with_st);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

int project_mg_dev(const tbfg_grid_desc* g, double alpha_dt, const double* rho, const double* psi, double* ux,
                   double* uy, double* uz, void* stream)
{
    if (!rho || !psi || !ux || !uy || !uz) {
        set_error("tbfg_project_mg_dev: NULL argument");
        return TBFG_EINVAL;
    }
    FieldGeo G;
    TBFG_TRY(get_geo(g, G));
    dim3 blk = row_block(G.nx);
    hipLaunchKernelGGL(k_project_mg, dim3(ceil_div(G.nx, blk.x) * G.ny * G.nz), blk, 0, (hipStream_t)stream, G,
                       alpha_dt, rho, psi, ux, uy, uz);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

}  // namespace tbfg
