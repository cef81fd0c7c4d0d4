// vof.hip -- material properties from the volume fraction on the device
// (updateMaterialProps and smoothVF, VOF.f90:197-262): two passes of
// trilinear cell -> vertex interpolation (interpolation.f90:188-226) and
// vertex -> cell averaging (:231-276) smooth c into cs, then
// rho = rho_g cs + (1 - cs) rho_l and mu likewise.  Ghost cells of c and cs
// are the periodic images in x and z and the wall row itself in y (the
// zero-gradient boundary of the volume fraction).
//
// Both passes are gathers over a 2x2x2 neighbourhood: one thread per vertex,
// then one per cell, consecutive threads along x (coalesced rows).  The
// coordinate tables (centres 0..n+1, faces 0..n per axis) are staged per grid.
#include <map>
#include <mutex>
#include <string>

#include "fields.hpp"

namespace tbfg {

struct VofGeo {
    int nx, ny, nz;
    const double* c[3];  // centre coordinates, index i = 0..n+1
    const double* f[3];  // face coordinates, index i = 0..n
};

struct VofGeoOwner {
    DevBuf buf;
    VofGeo view{};
};

static std::mutex g_vof_mu;
static std::map<std::string, std::unique_ptr<VofGeoOwner>> g_vof_cache;

static int get_vof_geo(const tbfg_grid_desc* g, VofGeo& out)
{
    TBFG_TRY(validate_grid(g));
    int dev = 0;
    TBFG_HIP(hipGetDevice(&dev));
    const std::string key = grid_key(g, dev);
    std::lock_guard<std::mutex> lk(g_vof_mu);
    auto it = g_vof_cache.find(key);
    if (it == g_vof_cache.end()) {
        HostLevel L = make_level(g, 0);
        std::vector<double> h;
        size_t off[3][2];
        for (int a = 0; a < 3; ++a) {
            off[a][0] = h.size();
            for (int i = 0; i <= L.n[a] + 1; ++i) h.push_back(L.ax[a].c(i));
            off[a][1] = h.size();
            for (int i = 0; i <= L.n[a]; ++i) h.push_back(L.ax[a].f(i));
        }
        std::unique_ptr<VofGeoOwner> o(new VofGeoOwner());
        TBFG_TRY(upload(o->buf, h.data(), h.size() * 8));
        o->view.nx = L.n[0];
        o->view.ny = L.n[1];
        o->view.nz = L.n[2];
        for (int a = 0; a < 3; ++a) {
            o->view.c[a] = o->buf.as<double>() + off[a][0];
            o->view.f[a] = o->buf.as<double>() + off[a][1];
        }
        it = g_vof_cache.emplace(key, std::move(o)).first;
    }
    out = it->second->view;
    return TBFG_OK;
}

// the cell value of a 0-based index with the ghosts of the volume fraction
__device__ __forceinline__ double cell_g(const VofGeo& G, const double* q, int i, int j, int k)
{
    i = i < 0 ? i + G.nx : (i >= G.nx ? i - G.nx : i);
    k = k < 0 ? k + G.nz : (k >= G.nz ? k - G.nz : k);
    j = j < 0 ? 0 : (j >= G.ny ? G.ny - 1 : j);
    return q[i + (size_t)G.nx * (j + (size_t)G.ny * k)];
}

// cellToVertex: vertex (i, j, k), i = 0..nx (face coordinates), from the
// centres i, i+1 (1-based; 0-based cells i-1, i)
__global__ void k_cell_to_vertex(VofGeo G, const double* __restrict__ q, double* __restrict__ qv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, k = blockIdx.z;
    if (i > G.nx) return;
    const double *x = G.c[0], *y = G.c[1], *z = G.c[2];
    const double xv = G.f[0][i], yv = G.f[1][j], zv = G.f[2][k];
    const double ax = x[i + 1] - xv, bx = xv - x[i];
    const double ay = y[j + 1] - yv, by = yv - y[j];
    const double az = z[k + 1] - zv, bz = zv - z[k];
    const double w1 = ax * by * az, w2 = bx * by * az, w3 = ax * by * bz, w4 = bx * by * bz;
    const double w5 = ax * ay * az, w6 = bx * ay * az, w7 = ax * ay * bz, w8 = bx * ay * bz;
    const double V = (x[i + 1] - x[i]) * (y[j + 1] - y[j]) * (z[k + 1] - z[k]);
    // 1-based centre m is 0-based cell m-1
    const double v = w1 * cell_g(G, q, i - 1, j, k - 1) + w2 * cell_g(G, q, i, j, k - 1) +
                     w3 * cell_g(G, q, i - 1, j, k) + w4 * cell_g(G, q, i, j, k) +
                     w5 * cell_g(G, q, i - 1, j - 1, k - 1) + w6 * cell_g(G, q, i, j - 1, k - 1) +
                     w7 * cell_g(G, q, i - 1, j - 1, k) + w8 * cell_g(G, q, i, j - 1, k);
    qv[i + (size_t)(G.nx + 1) * (j + (size_t)(G.ny + 1) * k)] = v / V;
}

// vertexToCell: 0-based cell (i, j, k) = 1-based (i+1, j+1, k+1), from the
// vertices i..i+1 (faces); with props, the material properties of the result
__global__ void k_vertex_to_cell(VofGeo G, const double* __restrict__ qv, double* __restrict__ q, double rho_l,
                                 double rho_g, double mu_l, double mu_g, double* __restrict__ rho,
                                 double* __restrict__ mu)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, k = blockIdx.z;
    if (i >= G.nx) return;
    const double xc = G.c[0][i + 1], yc = G.c[1][j + 1], zc = G.c[2][k + 1];
    const double *xf = G.f[0], *yf = G.f[1], *zf = G.f[2];
    const double ax = xf[i + 1] - xc, bx = xc - xf[i];
    const double ay = yc - yf[j], by = yf[j + 1] - yc;
    const double az = zf[k + 1] - zc, bz = zc - zf[k];
    const double w1 = ax * ay * az, w2 = bx * ay * az, w3 = ax * ay * bz, w4 = bx * ay * bz;
    const double w5 = ax * by * az, w6 = bx * by * az, w7 = ax * by * bz, w8 = bx * by * bz;
    const double V = (xf[i + 1] - xf[i]) * (yf[j + 1] - yf[j]) * (zf[k + 1] - zf[k]);
    const size_t sx = 1, sy = (size_t)(G.nx + 1), sz = (size_t)(G.nx + 1) * (G.ny + 1);
    const size_t v0 = i + sy * j + sz * k;  // vertex (i, j, k): lower corner
    const double v = w1 * qv[v0 + sy] + w2 * qv[v0 + sx + sy] + w3 * qv[v0 + sy + sz] + w4 * qv[v0 + sx + sy + sz] +
                     w5 * qv[v0] + w6 * qv[v0 + sx] + w7 * qv[v0 + sz] + w8 * qv[v0 + sx + sz];
    const double cs = v / V;
    const size_t c = i + (size_t)G.nx * (j + (size_t)G.ny * k);
    q[c] = cs;
    if (rho) rho[c] = rho_g * cs + (1.0 - cs) * rho_l;
    if (mu) mu[c] = mu_g * cs + (1.0 - cs) * mu_l;
}

int64_t material_work_doubles(const tbfg_grid_desc* g)
{
    if (validate_grid(g) != TBFG_OK) return -1;
    return (int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1);
}

int material_props_dev(const tbfg_grid_desc* g, double rho_l, double rho_g, double mu_l, double mu_g, const double* c,
                       double* cs, double* rho, double* mu, double* work, void* stream)
{
    if (!c || !cs || !rho || !work) {
        set_error("tbfg_material_props_dev: NULL argument");
        return TBFG_EINVAL;
    }
    if (c == cs) {
        set_error("tbfg_material_props_dev: c and cs must not alias");
        return TBFG_EINVAL;
    }
    VofGeo G;
    TBFG_TRY(get_vof_geo(g, G));
    hipStream_t s = (hipStream_t)stream;
    const dim3 blk(256);
    const dim3 gv(ceil_div(G.nx + 1, 256), G.ny + 1, G.nz + 1), gc(ceil_div(G.nx, 256), G.ny, G.nz);
    hipLaunchKernelGGL(k_cell_to_vertex, gv, blk, 0, s, G, c, work);
    hipLaunchKernelGGL(k_vertex_to_cell, gc, blk, 0, s, G, work, cs, 0.0, 0.0, 0.0, 0.0, (double*)nullptr,
                       (double*)nullptr);
    hipLaunchKernelGGL(k_cell_to_vertex, gv, blk, 0, s, G, cs, work);
    hipLaunchKernelGGL(k_vertex_to_cell, gc, blk, 0, s, G, work, cs, rho_l, rho_g, mu_l, mu_g, rho, mu);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

}  // namespace tbfg
