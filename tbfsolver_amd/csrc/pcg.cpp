// pcg.cpp -- host driver of the PCG + multigrid solve.
//
// Reference flow (pcg.f90:103-140): r = q - A psi; d = M r; delta = r.d; loop
// { qv = A d; alpha = delta/(d.qv); psi += alpha d; r -= alpha qv; z = M r;
// gamma = r.z/delta; d = z + gamma d } until ||r||/||q|| <= tolPCG.  M is one
// V-cycle (multiGrid.f90:186-232) with red/black Gauss-Seidel smoothing and an
// exact solve on the coarsest level (poissMat.f90:289-312).
//
// Device design: every scalar of the iteration lives in a PcgState on the
// device; kernels read alpha/gamma from it and the last block of each
// reduction writes them, so an iteration is a chain of launches with one
// host synchronisation at its end (the stop test).  Levels with at most
// kTailCells cells, the reference's coarsest level and the sub-hierarchy that
// approximates its exact solve run inside one single-workgroup kernel.
#include "pcg.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <utility>

#include "mg_launch.hpp"

using namespace tbfg;

namespace {

size_t tail_cells_limit()
{
    const char* e = std::getenv("TBFG_TAIL_CELLS");
    if (e && *e) {
        long v = std::atol(e);
        if (v > 0) return (size_t)v;
    }
    return 4096;
}

constexpr int kMaxBottomCells = 512;
constexpr int kMaxLevels = 32;

}  // namespace

tbfg_pcg::~tbfg_pcg()
{
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    if (hst) (void)hipHostFree(hst);
    if (hring) (void)hipHostFree(hring);
    for (hipEvent_t e : hev)
        if (e) (void)hipEventDestroy(e);
    if (hstage) (void)hipHostFree(hstage);
    if (hin) (void)hipHostFree(hin);
    if (hmg) (void)hipHostFree(hmg);
    if (fps) tbfg::fps_destroy(fps);
}

namespace tbfg {

// --------------------------------------------------------- kernel timing

static constexpr int kProfPairs = 1024;

// records an event pair around a launch of the profiled kernel family
// ext: the launch records the pair itself around the kernel (events())
struct ProfScope {
    tbfg_pcg* h;
    hipStream_t s;
    bool on, ext;
    KernelEvents kev{};
    ProfScope(tbfg_pcg* h_, int kernel, void* stream, bool ext_ = false) : h(h_), s((hipStream_t)stream), ext(ext_)
    {
        on = kernel >= 0 && h->prof_kernel == kernel && h->prof_used < kProfPairs;
        if (on && ext) kev = KernelEvents{h->prof_ev[2 * h->prof_used], h->prof_ev[2 * h->prof_used + 1]};
        if (on && !ext) (void)hipEventRecord(h->prof_ev[2 * h->prof_used], s);
    }
    const KernelEvents* events() const { return on && ext ? &kev : nullptr; }
    ~ProfScope()
    {
        if (on) {
            if (!ext) (void)hipEventRecord(h->prof_ev[2 * h->prof_used + 1], s);
            h->prof_used++;
        }
    }
};

// adds up the recorded pairs (the launches of one solve); the stream must have
// been synchronised.  Launches queued after the solve's stop test return at
// once (their kernels read the stop flag first): a launch shorter than a
// quarter of the solve's median is one of them and is not counted.
static void prof_harvest(tbfg_pcg* h)
{
    std::vector<float> ms;
    ms.reserve(h->prof_used);
    for (int i = 0; i < h->prof_used; ++i) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, h->prof_ev[2 * i], h->prof_ev[2 * i + 1]) == hipSuccess) ms.push_back(t);
    }
    h->prof_used = 0;
    if (ms.empty()) return;
    std::vector<float> sorted(ms);
    std::nth_element(sorted.begin(), sorted.begin() + sorted.size() / 2, sorted.end());
    const float cut = 0.25f * sorted[sorted.size() / 2];
    for (float t : ms)
        if (t >= cut) {
            h->prof_ms += t;
            h->prof_launches++;
        }
}

int pcg_profile_begin(tbfg_pcg* h, int kernel)
{
    if (!h || kernel < 0 || kernel > 3) {
        set_error("tbfg_pcg_profile_begin: bad handle or kernel id %d", kernel);
        return TBFG_EINVAL;
    }
    if (h->prof_ev.empty()) {
        h->prof_ev.resize(2 * kProfPairs);
        for (auto& e : h->prof_ev) TBFG_HIP(hipEventCreate(&e));
    }
    h->prof_kernel = kernel;
    h->prof_used = 0;
    h->prof_launches = 0;
    h->prof_ms = 0.0;
    return TBFG_OK;
}

int pcg_profile_end(tbfg_pcg* h, int64_t* launches, double* total_ms, double* bytes_per_launch)
{
    if (!h || h->prof_kernel < 0) {
        set_error("tbfg_pcg_profile_end: profiling not started");
        return TBFG_EINVAL;
    }
    TBFG_HIP(hipDeviceSynchronize());
    prof_harvest(h);
    const double n = (double)h->n0;
    const double bp = h->mixed ? 4.0 : 8.0;  // MG storage of p (and beta, q) on the finest level
    const double bq = h->b16 ? 2.0 : bp;      // ... of beta and q
    const double rq = h->b16 ? 2.0 : 4.0;     // the multigrid's copy of r (mixed)
    double bytes = 0.0;
    switch (h->prof_kernel) {
        case 0: bytes = (2.0 * bp + 2.0 * bq) * n; break;                 // full sweep: p, beta, q in; p out
        case 1: bytes = (bp + 2.0 * bq) * n + bp * n / 8; break;          // p, beta, q in; coarse q out
        case 2:  // d, beta in; qv out (fused: z, d_old, beta, psi in; d, qv, psi out)
            bytes = h->fuse_op ? (48.0 + (h->mixed && !h->fps ? 4.0 : 8.0)) * n : 24.0 * n;
            break;
        case 3:  // psi, r, d, qv in; psi, r (+ FP32 r) out (fused: r, qv in; r (+ FP32 r) out)
            bytes = (h->fuse_op ? 24.0 * n : 48.0 * n) + (h->mixed && !h->fps ? rq * n : 0.0);
            break;
    }
    if (launches) *launches = h->prof_launches;
    if (total_ms) *total_ms = h->prof_ms;
    if (bytes_per_launch) *bytes_per_launch = bytes;
    h->prof_kernel = -1;
    return TBFG_OK;
}

template <class T>
static const LevelDev<T>& geo(tbfg_pcg* h, int l);
template <>
const LevelDev<double>& geo<double>(tbfg_pcg* h, int l) { return h->tabd[l]->view; }
template <>
const LevelDev<float>& geo<float>(tbfg_pcg* h, int l) { return h->tabf[l]->view; }

template <class T>
static int build_plan(tbfg_pcg* h)
{
    TailPlan<T> P;
    std::memset(&P, 0, sizeof P);
    P.nlev = h->ntot - h->t0;
    // with coarse_max_iter == 0 the V-cycle runs straight to the dense bottom
    // and so does it when the coarsest user level is solved outside the tail
    // (the tail is then the bottom part of that solve's inner V-cycle)
    P.user_last = (h->prm.coarse_max_iter > 0 && h->kin < 0 ? h->nuser - 1 : h->ntot - 1) - h->t0;
    P.rev_post = h->kin >= 0 ? 1 : 0;
    P.bottom = (h->ntot - 1) - h->t0;
    P.n_pre = h->prm.n_pre;
    P.n_post = h->prm.n_post;
    P.cmaxit = h->prm.coarse_max_iter;
    P.ctol = (T)h->prm.coarse_tol;
    P.cx = h->cx.as<T>();
    P.cr = h->cr.as<T>();
    P.cd = h->cd.as<T>();
    P.cad = h->cad.as<T>();
    P.cvol = h->cvol.as<T>();
    P.G = h->G.as<double>();
    P.st = h->st.as<PcgState>();
    // LDS layout of the LDS-resident tail: per level p, q, beta, then its
    // table block (LevelTables::build), padded to 4 entries
    size_t lds = 0;
    int first = 0;
    auto add_seg = [&](const T* src, size_t dst, size_t n) {
        P.seg_src[P.nseg] = src;
        P.seg_dst[P.nseg] = (int)dst;
        P.seg_first[P.nseg] = first;
        first += (int)n;
        ++P.nseg;
    };
    for (int m = 0; m < P.nlev; ++m) {
        const int l = h->t0 + m;
        const size_t n = h->hl[l].cells();
        P.lv[m].geo = geo<T>(h, l);
        P.lv[m].p = h->p[l].as<T>();
        P.lv[m].q = h->q[l].as<T>();
        P.lv[m].beta = h->beta[l].as<T>();
        const size_t e0 = lds / sizeof(T);
        P.lv[m].lds_p = (int)e0;
        P.lv[m].lds_q = (int)(e0 + n);
        P.lv[m].lds_beta = (int)(e0 + 2 * n);
        P.lv[m].lds_tab = (int)(e0 + 3 * n);
        auto magic = [](int d) {
            return d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d);
        };
        if ((uint64_t)n * (uint64_t)std::max(h->hl[l].n[0], h->hl[l].n[1]) >= ((uint64_t)1 << 32)) {
            set_error("tail level %d x %d x %d too large for its index arithmetic", h->hl[l].n[0], h->hl[l].n[1],
                      h->hl[l].n[2]);
            return TBFG_EINVAL;
        }
        P.lv[m].dm_nx = magic(h->hl[l].n[0]);
        P.lv[m].dm_ny = magic(h->hl[l].n[1]);
        P.lv[m].dm_hx = magic(h->hl[l].n[0] / 2);
        P.lv[m].dm_hy = magic(h->hl[l].n[1] / 2);
        lds += 3 * n * sizeof(T);
        size_t nt = 0;
        for (int a = 0; a < 3; ++a)
            nt += 4 * (size_t)h->hl[l].n[a] + (m + 1 < P.nlev ? 2 * (size_t)h->hl[l + 1].n[a] : 0);
        lds += ((nt + 3) & ~(size_t)3) * sizeof(T);
        add_seg(P.lv[m].beta, P.lv[m].lds_beta, n);
        add_seg(P.lv[m].geo.a[0].ip, P.lv[m].lds_tab, nt);
    }
    add_seg(P.lv[0].q, P.lv[0].lds_q, h->hl[h->t0].cells());
    P.seg_first[P.nseg] = first;
    // the LDS-resident tail when its levels fit (the inner coarsest-level CG
    // keeps its vectors in HBM and uses the other form)
    h->tail_lds = 0;
    if ((h->prm.coarse_max_iter == 0 || h->kin >= 0) && !std::getenv("TBFG_TAIL_HBM") && lds <= tail_lds_capacity<T>())
        h->tail_lds = lds;
    return upload(h->plan, &P, sizeof P);
}

int pcg_create(const tbfg_grid_desc* g, const tbfg_pcg_params* prm_in, tbfg_pcg** out)
{
    if (!out || !prm_in) {
        set_error("tbfg_pcg_create: NULL argument");
        return TBFG_EINVAL;
    }
    *out = nullptr;
    TBFG_TRY(validate_grid(g));
    tbfg_pcg_params prm = *prm_in;
    if (prm.n_levels < 2 || prm.n_levels > 12) {
        set_error("n_levels = %d (expected 2..12)", prm.n_levels);
        return TBFG_EINVAL;
    }
    if (prm.n_pre < 0 || prm.n_post < 0 || prm.n_pre + prm.n_post < 1 || prm.mg_cycles < 1 || prm.max_iter < 0 ||
        !(prm.tol >= 0.0) || prm.coarse_max_iter < -1 || !(prm.coarse_tol >= 0.0) || !(prm.mg_tol >= 0.0) ||
        !(prm.mg_tol < 1.0) || (prm.mg_tol > 0.0 && prm.mg_max_iter < 1)) {
        set_error("invalid PCG/MG parameters (pre %d post %d cycles %d max_iter %d tol %g coarse %d/%g)", prm.n_pre,
                  prm.n_post, prm.mg_cycles, prm.max_iter, prm.tol, prm.coarse_max_iter, prm.coarse_tol);
        return TBFG_EINVAL;
    }
    if (prm.precision != TBFG_PREC_FP64 && prm.precision != TBFG_PREC_MIXED && prm.precision != TBFG_PREC_MIXED_BF16) {
        set_error("precision = %d", prm.precision);
        return TBFG_EINVAL;
    }
    if (prm.preconditioner != TBFG_PRECOND_MG && prm.preconditioner != TBFG_PRECOND_FFT) {
        set_error("preconditioner = %d", prm.preconditioner);
        return TBFG_EINVAL;
    }
    // multiGrid.f90:116-145: every axis divisible by 2^(nLevels-1)
    for (int a = 0; a < 3; ++a)
        if (g->n[a] % (1 << (prm.n_levels - 1)) != 0) {
            set_error("Attempt to initialize a MG grid with a non-integer number of cells along axis %d "
                      "(n=%d, levels=%d)",
                      a, g->n[a], prm.n_levels);
            return TBFG_EINVAL;
        }
    std::unique_ptr<tbfg_pcg> h(new tbfg_pcg());
    h->g = *g;
    h->prm = prm;
    h->mixed = prm.precision == TBFG_PREC_MIXED || prm.precision == TBFG_PREC_MIXED_BF16;
    if (h->prm.flexible < 0) h->prm.flexible = h->mixed ? 1 : 0;
    if (!(h->prm.coarse_tol > 0.0)) h->prm.coarse_tol = h->mixed ? 1e-5 : 1e-10;
    if (h->prm.coarse_max_iter < 0) h->prm.coarse_max_iter = h->mixed ? 0 : 40;
    prm.coarse_max_iter = h->prm.coarse_max_iter;
    h->nuser = prm.n_levels;
    // sub-hierarchy under the coarsest level: halve while every axis is even
    // The sub-hierarchy serves the coarsest-level solve: with coarse_max_iter > 0
    // as the preconditioner of an inner CG, with coarse_max_iter == 0 as the
    // continuation of the V-cycle itself (one cycle from zero, dense bottom).
    int ntot = h->nuser;
    {
        for (;;) {
            bool ok = true;
            size_t cells = 1;
            for (int a = 0; a < 3; ++a) {
                int nn = g->n[a] >> (ntot - 1);
                cells *= (size_t)nn;
                if (nn % 2 != 0 || nn < 2) ok = false;
            }
            if (!ok || cells <= 64 || ntot >= 20) break;
            ntot++;
        }
    }
    h->ntot = ntot;
    {
        size_t cb = 1;
        for (int a = 0; a < 3; ++a) cb *= (size_t)(g->n[a] >> (ntot - 1));
        if (cb > (size_t)kMaxBottomCells) {
            set_error("the coarsest grid has %zu cells after halving every axis while all are even; the dense "
                      "bottom solve needs <= %d",
                      cb, kMaxBottomCells);
            return TBFG_EINVAL;
        }
    }
    const size_t tail_lim = tail_cells_limit();
    // first level of the single-workgroup tail: the first small level, and at
    // the latest the reference's coarsest level when that one is solved by the
    // inner CG (which then runs inside the tail).  A coarsest user level too
    // large for the tail has its inner CG run as multi-workgroup kernels, and
    // the tail starts at the first small level under it.
    auto level_cells = [&](int l) {
        size_t cells = 1;
        for (int a = 0; a < 3; ++a) cells *= (size_t)(g->n[a] >> l);
        return cells;
    };
    // (above ~2K cells the grid-wide kernels are the quicker form: 149 against
    // 162 ms per FP64 solve with a 16x12x16 coarsest level at 256x192x256)
    h->kin = -1;
    const size_t inner_lim = std::min<size_t>(tail_lim, 2048);
    if (prm.coarse_max_iter > 0 && ntot > h->nuser && level_cells(h->nuser - 1) > inner_lim &&
        !std::getenv("TBFG_INNER_IN_TAIL"))
        h->kin = h->nuser - 1;
    const int t_cap = prm.coarse_max_iter > 0 && h->kin < 0 ? h->nuser - 1 : ntot - 1;
    int t0 = t_cap;
    for (int l = h->kin >= 0 ? h->kin + 1 : 1; l < t_cap; ++l) {
        size_t cells = 1;
        for (int a = 0; a < 3; ++a) cells *= (size_t)(g->n[a] >> l);
        if (cells <= tail_lim) {
            t0 = l;
            break;
        }
    }
    h->t0 = t0;
    if (h->ntot - h->t0 > kMaxTail) {
        set_error("hierarchy too deep for the tail kernel (%d levels)", h->ntot - h->t0);
        return TBFG_EINVAL;
    }
    for (int l = 0; l < ntot; ++l) h->hl.push_back(make_level(g, l));
    h->n0 = h->hl[0].cells();
    // tables
    h->tabd.resize(ntot);
    for (int l = 0; l < ntot; ++l) {
        h->tabd[l].reset(new LevelTables<double>());
        TBFG_TRY(h->tabd[l]->build(h->hl[l], l + 1 < ntot ? &h->hl[l + 1] : nullptr));
    }
    if (h->mixed) {
        h->tabf.resize(ntot);
        for (int l = 0; l < ntot; ++l) {
            h->tabf[l].reset(new LevelTables<float>());
            TBFG_TRY(h->tabf[l]->build(h->hl[l], l + 1 < ntot ? &h->hl[l + 1] : nullptr));
        }
    }
    const size_t es = h->mixed ? sizeof(float) : sizeof(double);
    h->p.resize(ntot);
    h->palt.resize(ntot);
    h->q.resize(ntot);
    h->beta.resize(ntot);
    h->fused.assign(ntot, 0);
    for (int l = 0; l < ntot; ++l) {
        const size_t n = h->hl[l].cells();
        TBFG_TRY(h->p[l].alloc(n * es));
        if (l > 0) TBFG_TRY(h->q[l].alloc(n * es));
        if (l < (h->kin >= 0 ? h->kin : t0)) {
            h->fused[l] = h->mixed ? sweep_supported<float>(h->tabf[l]->view) : sweep_supported<double>(h->tabd[l]->view);
            // below ~1M cells a plane march of a few steps per workgroup is
            // latency bound and the two half-sweep launches are quicker
            size_t min_cells = (size_t)1 << 20;
            if (const char* e = std::getenv("TBFG_SWEEP_MIN_CELLS")) min_cells = (size_t)std::atoll(e);
            if (n < min_cells || std::getenv("TBFG_HALF_SWEEPS")) h->fused[l] = 0;
            // below that, one launch per sweep from bricks staged in LDS
            if (!h->fused[l] && !std::getenv("TBFG_HALF_SWEEPS") && !std::getenv("TBFG_NO_BRICKS") &&
                (h->mixed ? brick_supported<float>(h->tabf[l]->view) : brick_supported<double>(h->tabd[l]->view)))
                h->fused[l] = 2;
            if (h->fused[l]) TBFG_TRY(h->palt[l].alloc(n * es));
        }
        // bf16 storage of the finest level's q and beta: where the plane march smooths it
        if (l == 0)
            h->b16 = prm.precision == TBFG_PREC_MIXED_BF16 && h->fused[0] == 1 &&
                     prm.preconditioner != TBFG_PRECOND_FFT;
        if (l > 0 || h->mixed) TBFG_TRY(h->beta[l].alloc(n * (l == 0 && h->b16 ? sizeof(bf16s) : es)));
    }
    if (h->mixed) TBFG_TRY(h->r32.alloc(h->n0 * (h->b16 ? sizeof(bf16s) : sizeof(float))));
    h->fuse_dots = !std::getenv("TBFG_UNFUSED_DOTS");
    if (prm.preconditioner == TBFG_PRECOND_FFT) {
        // FP32 transforms without refinement in mixed precision (a preconditioner)
        tbfg_fps_opts fo{h->mixed ? TBFG_PREC_MIXED : TBFG_PREC_FP64, 0};
        TBFG_TRY(fps_create(g, &fo, &h->fps));
        TBFG_TRY(h->zfft.alloc(h->n0 * 8));
    }
    TBFG_TRY(h->r.alloc(h->n0 * 8));
    TBFG_TRY(h->d.alloc(h->n0 * 8));
    TBFG_TRY(h->qv.alloc(h->n0 * 8));
    TBFG_TRY(h->st.alloc(sizeof(PcgState)));
    h->opp = op_plan(h->tabd[0]->view);
    h->fuse_op = h->opp.tx != 0 && !std::getenv("TBFG_UNFUSED_OP");
    if (h->fuse_op) TBFG_TRY(h->d2.alloc(h->n0 * 8));
    TBFG_TRY(h->partials.alloc(reduction_slots(h->tabd[0]->view, h->opp) * sizeof(double)));
    TBFG_TRY(h->counter.alloc(256));
    TBFG_TRY(memset_sync(h->counter.p, 0, 256));
    // dense bottom operator
    const HostLevel& B = h->hl[ntot - 1];
    const size_t nbc = B.cells();
    TBFG_TRY(h->G.alloc(nbc * nbc * 8));
    TBFG_TRY(h->dwork.alloc((2 * nbc * nbc + nbc + 16) * 8));
    {
        std::vector<double> vol(nbc);
        for (int k = 1; k <= B.n[2]; ++k)
            for (int j = 1; j <= B.n[1]; ++j)
                for (int i = 1; i <= B.n[0]; ++i)
                    vol[(i - 1) + (size_t)B.n[0] * ((j - 1) + (size_t)B.n[1] * (k - 1))] =
                        B.ax[0].df(i) * B.ax[1].df(j) * B.ax[2].df(k);
        TBFG_TRY(upload(h->vol, vol.data(), nbc * 8));
    }
    // inner CG vectors and cell volumes of the coarsest user level
    if (prm.coarse_max_iter > 0) {
        const HostLevel& U = h->hl[h->nuser - 1];
        const size_t nu = U.cells();
        TBFG_TRY(h->cx.alloc(nu * es));
        TBFG_TRY(h->cr.alloc(nu * es));
        TBFG_TRY(h->cd.alloc(nu * es));
        TBFG_TRY(h->cad.alloc(nu * es));
        if (h->kin >= 0) {
            TBFG_TRY(h->instate.alloc(sizeof(InnerState)));
            TBFG_HIP(hipHostMalloc((void**)&h->hin, 2 * sizeof(int32_t), hipHostMallocDefault));
        }
        std::vector<double> v(nu);
        for (int k = 1; k <= U.n[2]; ++k)
            for (int j = 1; j <= U.n[1]; ++j)
                for (int i = 1; i <= U.n[0]; ++i)
                    v[(i - 1) + (size_t)U.n[0] * ((j - 1) + (size_t)U.n[1] * (k - 1))] =
                        U.ax[0].df(i) * U.ax[1].df(j) * U.ax[2].df(k);
        if (h->mixed) {
            std::vector<float> vf(v.begin(), v.end());
            TBFG_TRY(upload(h->cvol, vf.data(), nu * 4));
        } else {
            TBFG_TRY(upload(h->cvol, v.data(), nu * 8));
        }
    }
    if (h->mixed)
        TBFG_TRY(build_plan<float>(h.get()));
    else
        TBFG_TRY(build_plan<double>(h.get()));
    TBFG_HIP(hipHostMalloc((void**)&h->hst, sizeof(PcgState), hipHostMallocDefault));
    std::memset(h->hst, 0, sizeof(PcgState));
    TBFG_HIP(hipHostMalloc((void**)&h->hring, 2 * sizeof(PcgState), hipHostMallocDefault));
    TBFG_HIP(hipHostMalloc((void**)&h->hmg, sizeof(int32_t), hipHostMallocDefault));
    for (auto& e : h->hev) TBFG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // footprint: every device buffer of the handle
    int64_t fp = 0;
    for (auto& b : h->p) fp += (int64_t)b.bytes;
    for (auto& b : h->palt) fp += (int64_t)b.bytes;
    for (auto& b : h->q) fp += (int64_t)b.bytes;
    for (auto& b : h->beta) fp += (int64_t)b.bytes;
    for (auto& t : h->tabd) fp += (int64_t)t->buf.bytes;
    for (auto& t : h->tabf) fp += (int64_t)t->buf.bytes;
    fp += (int64_t)(h->r.bytes + h->r32.bytes + h->d.bytes + h->qv.bytes + h->st.bytes + h->partials.bytes + h->counter.bytes +
                    h->plan.bytes + h->G.bytes + h->dwork.bytes + h->vol.bytes + h->cx.bytes + h->cr.bytes +
                    h->cd.bytes + h->cad.bytes + h->cvol.bytes + h->zfft.bytes + h->d2.bytes +
                    h->instate.bytes);
    if (h->fps) fp += fps_footprint(h->fps);
    h->footprint = fp;
    *out = h.release();
    return TBFG_OK;
}

// beta on every level (restricted once per solve; multiGrid.f90:210 restricts it
// every V-cycle, but it does not change within a solve) and the dense bottom.
template <class T, class T0 = T>
static int setup_levels(tbfg_pcg* h, const double* beta0, void* stream)
{
    const T* prev;
    int l = 1;
    if constexpr (std::is_same<T0, bf16s>::value) {
        // bf16 finest level: level 1 restricted from the FP64 beta
        TBFG_TRY(launch_convert<double, bf16s>(h->n0, beta0, h->beta[0].as<bf16s>(), stream));
        TBFG_TRY(launch_restrict<double, float>(geo<float>(h, 0), beta0, h->beta[1].as<float>(), stream));
        prev = h->beta[1].as<T>();
        l = 2;
    } else if constexpr (sizeof(T) == 4) {
        TBFG_TRY(launch_convert<double, float>(h->n0, beta0, h->beta[0].as<float>(), stream));
        prev = h->beta[0].as<T>();
    } else {
        prev = beta0;
    }
    for (; l < h->ntot; ++l) {
        TBFG_TRY(launch_restrict<T, T>(geo<T>(h, l - 1), prev, h->beta[l].as<T>(), stream));
        prev = h->beta[l].as<T>();
    }
    TBFG_TRY(launch_dense_setup<T>(h->tabd[h->ntot - 1]->view, prev, h->vol.as<double>(), h->dwork.as<double>(),
                                   h->G.as<double>(), stream));
    return TBFG_OK;
}

// nsweeps red/black sweeps on level l (rbgs.f90:205-259); `first`: the level
// starts from p = 0.  Levels with the full-sweep kernel alternate between
// their two buffers, so p (and alt) may be swapped on return.
// last_dots: the dot products of the PCG are formed by the last sweep (see
// launch_sweep); *nb_out receives its partial count.
// (T0: the storage type of the level's q and beta, T or bf16s; bf16s only
// where the plane march smooths the level)
template <class T, class T0 = T>
static int smooth(tbfg_pcg* h, int l, const T0* b, const T0* q, T*& p, T*& alt, int nsweeps, bool first,
                  const int* done, void* stream, const SweepDots* last_dots = nullptr, unsigned* nb_out = nullptr)
{
    const LevelDev<T>& L = geo<T>(h, l);
    constexpr bool narrow = !std::is_same<T0, T>::value;
    if (narrow && h->fused[l] != 1) {
        set_error("smooth: bf16 storage on a level without the plane-march sweep");
        return TBFG_EINVAL;
    }
    for (int s = 0; s < nsweeps; ++s) {
        const bool f0 = first && s == 0;
        if constexpr (narrow) {
            const bool last = s == nsweeps - 1;
            const bool plain = l == 0 && !f0 && !(last && last_dots);
            ProfScope ps(h, plain ? 0 : -1, stream, true);
            TBFG_TRY(launch_sweep<T, T0, T0>(L, b, q, p, alt, f0, done, stream, last ? last_dots : nullptr,
                                             last ? nb_out : nullptr, ps.events()));
            std::swap(p, alt);
        } else if (h->fused[l] == 2) {
            ProfScope ps(h, l == 0 ? 0 : -1, stream);
            TBFG_TRY(launch_brick_sweep<T, T>(L, b, q, p, alt, f0, done, stream));
            std::swap(p, alt);
        } else if (h->fused[l]) {
            const bool last = s == nsweeps - 1;
            // the timed finest-level sweep: from a non-zero start, without the dot products
            const bool plain = l == 0 && !f0 && !(last && last_dots);
            ProfScope ps(h, plain ? 0 : -1, stream, true);
            TBFG_TRY(launch_sweep<T, T>(L, b, q, p, alt, f0, done, stream, last ? last_dots : nullptr,
                                        last ? nb_out : nullptr, ps.events()));
            std::swap(p, alt);
        } else {
            ProfScope ps(h, l == 0 ? 0 : -1, stream);
            TBFG_TRY(launch_rbgs<T, T>(L, b, q, p, RED, f0, done, stream));
            TBFG_TRY(launch_rbgs<T, T>(L, b, q, p, BLACK, f0, done, stream));
        }
    }
    return TBFG_OK;
}

// The inner V-cycle of the coarsest-level solve on level kin (the tail's
// tail_inner_vcycle, for a level too large for the tail): from zero, red/black
// pre-smoothing and black/red post-smoothing in place on the levels above the
// tail, the tail (post-smoothing reversed too) below.  q[kin] in, p[kin] out.
template <class T>
static int inner_vcycle(tbfg_pcg* h, const int* done, void* stream)
{
    const int t0 = h->t0;
    for (int l = h->kin; l < t0; ++l) {
        const LevelDev<T>& L = geo<T>(h, l);
        const T* b = h->beta[l].as<T>();
        const T* q = h->q[l].as<T>();
        T* p = h->p[l].as<T>();
        if (h->prm.n_pre == 0)
            TBFG_HIP(hipMemsetAsync(p, 0, h->hl[l].cells() * sizeof(T), (hipStream_t)stream));
        for (int s = 0; s < h->prm.n_pre; ++s) {
            TBFG_TRY(launch_rbgs<T, T>(L, b, q, p, RED, s == 0, done, stream));
            TBFG_TRY(launch_rbgs<T, T>(L, b, q, p, BLACK, s == 0, done, stream));
        }
        TBFG_TRY(launch_resid_restrict<T, T>(L, b, q, p, h->q[l + 1].as<T>(), done, stream));
    }
    TBFG_TRY(launch_tail<T>(h->plan.as<TailPlan<T>>(), done, h->tail_lds, stream));
    for (int l = t0 - 1; l >= h->kin; --l) {
        const LevelDev<T>& L = geo<T>(h, l);
        const T* b = h->beta[l].as<T>();
        const T* q = h->q[l].as<T>();
        T* p = h->p[l].as<T>();
        TBFG_TRY(launch_prolong_add<T>(L, h->p[l + 1].as<T>(), p, done, stream));
        for (int s = 0; s < h->prm.n_post; ++s) {
            TBFG_TRY(launch_rbgs<T, T>(L, b, q, p, BLACK, false, done, stream));
            TBFG_TRY(launch_rbgs<T, T>(L, b, q, p, RED, false, done, stream));
        }
    }
    return TBFG_OK;
}

// The coarsest-level solve (poissMat.f90:289-312) as the tail's
// tail_coarse_pcg computes it, by multi-workgroup kernels: flexible CG in the
// V-weighted inner product preconditioned by inner_vcycle, to the relative
// residual coarse_tol or coarse_max_iter iterations.  q[kin] in, p[kin] out.
// The host queues as many iterations as the previous solve took, then reads
// the stop flag and queues two more at a time until it is set; iterations
// queued after the stop return at once.
template <class T>
static int inner_solve(tbfg_pcg* h, void* stream)
{
    const int U = h->kin;
    const LevelDev<T>& L = geo<T>(h, U);
    InnerState* is = h->instate.as<InnerState>();
    PcgState* st = h->st.as<PcgState>();
    const int* idone = &is->done;
    Reducer red{h->partials.as<double>()};
    InnerArgs<T> a;
    a.beta = h->beta[U].as<T>();
    a.vol = h->cvol.as<T>();
    a.q = h->q[U].as<T>();
    a.r = h->cr.as<T>();
    a.x = h->cx.as<T>();
    a.d = h->cd.as<T>();
    a.ad = h->cad.as<T>();
    a.z = h->p[U].as<T>();
    a.p = h->p[U].as<T>();
    const double ctol = (double)(T)h->prm.coarse_tol;
    a.tol2 = ctol * ctol;
    TBFG_TRY(launch_inner<T>(InnerOp::Start, L, a, is, st, red, stream));
    TBFG_TRY(inner_vcycle<T>(h, idone, stream));
    TBFG_TRY(launch_inner<T>(InnerOp::FirstDir, L, a, is, st, red, stream));
    const int maxit = h->prm.coarse_max_iter;
    int queued = 0, chunk = std::max(1, h->inner_guess);
    while (queued < maxit) {
        const int m = std::min(chunk, maxit - queued);
        for (int it = 0; it < m; ++it) {
            TBFG_TRY(launch_inner<T>(InnerOp::OpUpdate, L, a, is, st, red, stream));
            TBFG_TRY(inner_vcycle<T>(h, idone, stream));
            TBFG_TRY(launch_inner<T>(InnerOp::NextDir, L, a, is, st, red, stream));
        }
        queued += m;
        if (queued >= maxit) break;
        TBFG_HIP(hipMemcpyAsync(h->hin, &is->done, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
        TBFG_HIP(hipStreamSynchronize((hipStream_t)stream));
        if (h->hin[0]) {
            if (h->hin[1] > 0) h->inner_guess = h->hin[1];
            break;
        }
        chunk = 2;
    }
    return launch_inner<T>(InnerOp::Finish, L, a, is, st, red, stream);
}

// One preconditioner application z = M r (multiGrid.f90:165-232, fixed cycles).
// r0 is the finest right-hand side in the multigrid precision; on return *z
// points at the level-0 buffer that holds the result.  dots_init >= 0: the
// finest level's last sweep also forms the PCG dot products of z (k_pcg_dots
// with that init flag, finalized here); *dots_done tells whether it could.
template <class T, class T0 = T>
static int vcycle(tbfg_pcg* h, const T0* r0, T** z, const T0* beta0, const int* done, void* stream,
                  int dots_init = -1, bool* dots_done = nullptr)
{
    if (dots_done) *dots_done = false;
    // solveMG's loop (mg_tol > 0): after every cycle the finest residual is
    // tested on the device and the host reads the decision before it queues
    // another cycle; the last cycle is not known in advance, so the dot
    // products are not folded into a sweep
    const bool tol_mode = h->prm.mg_tol > 0.0;
    const int ncycles = tol_mode ? h->prm.mg_max_iter : h->prm.mg_cycles;
    const bool want_dots = !tol_mode && dots_init >= 0 && h->fused[0] == 1 && h->prm.n_post > 0 && h->t0 > 0;
    SweepDots sd;
    if (want_dots) {
        sd.r = h->r.as<double>();
        sd.qv = h->qv.as<double>();
        sd.init = dots_init;
        sd.red = Reducer{h->partials.as<double>()};
        sd.capacity = h->partials.bytes / sizeof(double);
    }
    // levels smoothed here; below them the tail, or the multi-workgroup solve of level kin
    const int t0 = h->kin >= 0 ? h->kin : h->t0;
    T* cur[kMaxLevels];
    T* alt[kMaxLevels];
    for (int l = 0; l < t0; ++l) {
        cur[l] = h->p[l].as<T>();
        alt[l] = h->palt[l].as<T>();
    }
    for (int cyc = 0; cyc < ncycles; ++cyc) {
        for (int l = 0; l < t0; ++l) {
            const bool first = (l > 0) || (cyc == 0);
            if (h->prm.n_pre == 0 && first)
                TBFG_HIP(hipMemsetAsync(cur[l], 0, h->hl[l].cells() * sizeof(T), (hipStream_t)stream));
            if (l == 0) {
                TBFG_TRY(smooth<T, T0>(h, 0, beta0, r0, cur[0], alt[0], h->prm.n_pre, first, done, stream));
                ProfScope ps(h, 1, stream);
                TBFG_TRY(launch_resid_restrict<T, T0, T0>(geo<T>(h, 0), beta0, r0, cur[0], h->q[1].as<T>(), done,
                                                          stream));
            } else {
                const T* b = h->beta[l].as<T>();
                const T* q = h->q[l].as<T>();
                TBFG_TRY(smooth<T>(h, l, b, q, cur[l], alt[l], h->prm.n_pre, first, done, stream));
                TBFG_TRY(launch_resid_restrict<T, T>(geo<T>(h, l), b, q, cur[l], h->q[l + 1].as<T>(), done, stream));
            }
        }
        if (h->kin >= 0)
            TBFG_TRY(inner_solve<T>(h, stream));
        else
            TBFG_TRY(launch_tail<T>(h->plan.as<TailPlan<T>>(), done, h->tail_lds, stream));
        for (int l = t0 - 1; l >= 0; --l) {
            const T* pc = (l + 1 < t0) ? cur[l + 1] : h->p[l + 1].as<T>();
            TBFG_TRY(launch_prolong_add<T>(geo<T>(h, l), pc, cur[l], done, stream));
            const bool fuse = want_dots && l == 0 && cyc == ncycles - 1;
            unsigned nb = 0;
            if (l == 0)
                TBFG_TRY(smooth<T, T0>(h, 0, beta0, r0, cur[0], alt[0], h->prm.n_post, false, done, stream,
                                       fuse ? &sd : nullptr, fuse ? &nb : nullptr));
            else
                TBFG_TRY(smooth<T>(h, l, h->beta[l].as<T>(), h->q[l].as<T>(), cur[l], alt[l], h->prm.n_post, false,
                                   done, stream));
            if (fuse) {
                TBFG_TRY(launch_fin_dots(nb, h->n0, dots_init, h->st.as<PcgState>(), sd.red, stream));
                if (dots_done) *dots_done = true;
            }
        }
        if (tol_mode) {
            PcgState* st = h->st.as<PcgState>();
            TBFG_TRY(launch_mg_res<T, T0>(geo<T>(h, 0), beta0, r0, cur[0], h->prm.mg_tol, cyc, ncycles, st,
                                      Reducer{h->partials.as<double>()}, stream));
            if (cyc + 1 < ncycles) {
                TBFG_HIP(hipMemcpyAsync(h->hmg, &st->mg_stop, sizeof(int32_t), hipMemcpyDeviceToHost,
                                        (hipStream_t)stream));
                TBFG_HIP(hipStreamSynchronize((hipStream_t)stream));
                if (*h->hmg) break;
            }
        }
    }
    *z = cur[0];
    return TBFG_OK;
}

// The outcome of the last solve from the host copy of its state (iterations,
// residual, status; the cycle counts of fixed-cycle applications).
static int finish_state(tbfg_pcg* h, bool fft, int32_t* iters, double* res)
{
    if (fft) {
        h->hst->mg_cyc_total = h->hst->mg_cyc_max = 0;
    } else if (!(h->prm.mg_tol > 0.0)) {
        // fixed cycles: the first application and one after every update but
        // the last (whose stop test comes before it; the reference applies
        // the preconditioner once more before testing, pcg.f90:122-136)
        h->hst->mg_cyc_total = h->hst->iter * h->prm.mg_cycles;
        h->hst->mg_cyc_max = h->hst->iter > 0 ? h->prm.mg_cycles : 0;
    }
    if (iters) *iters = h->hst->iter;
    if (res) *res = h->hst->res;
    if (h->hst->status == TBFG_ENUMERIC) {
        set_error("PCG: non-finite value (res=%g) after %d iterations", h->hst->res, h->hst->iter);
        return TBFG_ENUMERIC;
    }
    if (h->hst->status == TBFG_ENOTCONVERGED) {
        set_error("EXIT PCG iterations: max iter reached (%d, res=%g)", h->hst->iter, h->hst->res);
        return TBFG_ENOTCONVERGED;
    }
    if (!h->hst->done) {
        set_error("PCG: not converged within the %d iterations queued (res=%g)", h->hst->iter, h->hst->res);
        return TBFG_ENOTCONVERGED;
    }
    return TBFG_OK;
}

// T: the multigrid's precision; FFT: the preconditioner is the FFT solve of
// the constant-coefficient operator (z in FP64) instead of the V-cycle; T0:
// the storage type of the finest level's multigrid q and beta (T or bf16s).
// queue > 0: enqueue exactly that many iterations and return at once, with no
// host synchronisation and no host memory in the stream (a solve a caller can
// capture in a graph); the device's stop test still ends the solve (iterations
// queued after it return at once) and pcg_result reads the outcome.
template <class T, bool FFT = false, class T0 = T>
static int solve_t(tbfg_pcg* h, double* psi, const double* beta, const double* q, int32_t* iters, double* res,
                   void* stream, int32_t queue = 0)
{
    using TZ = typename std::conditional<FFT, double, T>::type;
    hipStream_t s = (hipStream_t)stream;
    PcgState* st = h->st.as<PcgState>();
    const int* done = &st->done;
    Reducer red{h->partials.as<double>()};
    if (queue > 0 && (h->prm.mg_tol > 0.0 || h->kin >= 0 || h->prof_kernel >= 0 || h->prm.n_pre == 0)) {
        set_error("tbfg_pcg_solve_queued_dev: this handle's solve synchronises with the host (tolMG loop, "
                  "coarsest level outside the tail kernel, no pre-smoothing, or kernel profiling)");
        return TBFG_EINVAL;
    }
    TBFG_TRY(launch_state_init(st, h->prm.tol, h->prm.max_iter, h->prm.flexible, stream));
    if (!FFT) TBFG_TRY(setup_levels<T, T0>(h, beta, stream));
    const T0* beta0;
    if constexpr (sizeof(T) == 4)
        beta0 = h->beta[0].as<T0>();
    else
        beta0 = beta;
    TZ* z = nullptr;
    const LevelDev<double>& L0 = h->tabd[0]->view;
    const size_t n = h->n0;
    double* r = h->r.as<double>();
    double* d = h->d.as<double>();
    double* qv = h->qv.as<double>();
    using TR = typename std::conditional<std::is_same<T0, bf16s>::value, bf16s, float>::type;
    TR* r32 = (h->mixed && !FFT) ? h->r32.as<TR>() : nullptr;
    const T0* r0;
    if constexpr (sizeof(T) == 4)
        r0 = r32;
    else
        r0 = r;
    // z = M r and the dot products of z (pcg.f90:118-139, 294-327)
    bool dots_done = false;
    auto precondition = [&](int init) -> int {
        if constexpr (FFT) {
            z = h->zfft.as<double>();
            TBFG_TRY(fps_solve_dev(h->fps, r, z, stream));
            TBFG_TRY(launch_pcg_dots<double>(n, r, z, qv, init, st, red, stream));
        } else {
            TBFG_TRY(vcycle<T, T0>(h, r0, &z, beta0, done, stream, h->fuse_dots ? init : -1, &dots_done));
            if (!dots_done) TBFG_TRY(launch_pcg_dots<T>(n, r, z, qv, init, st, red, stream));
        }
        return TBFG_OK;
    };
    TBFG_TRY(launch_residual_init(L0, h->opp, beta, q, psi, r, r32, st, red, stream));
    TBFG_TRY(precondition(1));
    // one iteration (pcg.f90:142-290): d = (z - m) + gamma d, qv = A d, alpha,
    // psi += alpha d, r -= alpha qv, the stop test, z = M r, the dot products.
    // Fused form: iteration j writes its direction to dbuf[j & 1] from
    // dbuf[(j - 1) & 1] in the operator pass, which also adds the previous
    // iteration's alpha d to psi; the last one is added after the loop.
    double* dbuf[2] = {d, h->fuse_op ? h->d2.as<double>() : d};
    int j = 0;
    auto iteration = [&](int init) -> int {
        ++j;
        if (h->fuse_op) {
            double* dj = dbuf[j & 1];
            {
                ProfScope ps(h, 2, stream, true);  // the operator pass alone (not its finalize)
                TBFG_TRY(launch_op_dir<TZ>(L0, h->opp, beta, z, dbuf[(j - 1) & 1], dj, psi, init, qv, st, red,
                                           stream, ps.events()));
            }
            ProfScope ps(h, 3, stream);
            TBFG_TRY(launch_pcg_update(n, nullptr, r, r32, dj, qv, st, red, stream));
        } else {
            TBFG_TRY(launch_pcg_dir<TZ>(n, z, d, init, st, stream));
            {
                ProfScope ps(h, 2, stream);
                TBFG_TRY(launch_op_dot(L0, h->opp, beta, d, qv, st, red, stream));
            }
            ProfScope ps(h, 3, stream);
            TBFG_TRY(launch_pcg_update(n, psi, r, r32, d, qv, st, red, stream));
        }
        TBFG_TRY(precondition(0));
        return TBFG_OK;
    };
    // copy of the device state after the work enqueued so far, and its event
    auto snapshot = [&](int slot) -> int {
        TBFG_HIP(hipMemcpyAsync(&h->hring[slot], st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        TBFG_HIP(hipEventRecord(h->hev[slot], s));
        return TBFG_OK;
    };
    if (queue > 0) {
        for (int32_t it = 1; it <= queue; ++it) TBFG_TRY(iteration(it == 1 ? 1 : 0));
        if (h->fuse_op) TBFG_TRY(launch_psi_last(n, psi, dbuf[0], dbuf[1], st, stream));
        h->queued_fft = FFT;
        return TBFG_OK;
    }
    // The host examines the state after iteration i while iteration i+1 is
    // already queued, so the device does not idle through the stop test; the
    // kernels of an iteration queued after convergence return at once.
    TBFG_TRY(snapshot(0));
    TBFG_TRY(iteration(1));
    TBFG_TRY(snapshot(1));
    for (int slot = 0;; slot ^= 1) {
        TBFG_HIP(hipEventSynchronize(h->hev[slot]));
        if (h->hring[slot].done) break;
        TBFG_TRY(iteration(0));
        TBFG_TRY(snapshot(slot));
    }
    if (h->fuse_op) TBFG_TRY(launch_psi_last(n, psi, dbuf[0], dbuf[1], st, stream));
    TBFG_HIP(hipMemcpyAsync(h->hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    TBFG_HIP(hipStreamSynchronize(s));
    if (h->prof_kernel >= 0) prof_harvest(h);
    return finish_state(h, FFT, iters, res);
}

int pcg_solve_dev(tbfg_pcg* h, double* psi, const double* beta, const double* q, int32_t* iters, double* res,
                  void* stream)
{
    if (!h || !psi || !beta || !q) {
        set_error("tbfg_pcg_solve_dev: NULL argument");
        return TBFG_EINVAL;
    }
    if (h->fps)
        return solve_t<double, true>(h, psi, beta, q, iters, res, stream);
    if (h->b16) return solve_t<float, false, bf16s>(h, psi, beta, q, iters, res, stream);
    return h->mixed ? solve_t<float>(h, psi, beta, q, iters, res, stream)
                    : solve_t<double>(h, psi, beta, q, iters, res, stream);
}

int pcg_solve_queued_dev(tbfg_pcg* h, double* psi, const double* beta, const double* q, int32_t iterations,
                         void* stream)
{
    if (!h || !psi || !beta || !q) {
        set_error("tbfg_pcg_solve_queued_dev: NULL argument");
        return TBFG_EINVAL;
    }
    if (iterations < 1 || iterations > h->prm.max_iter) {
        set_error("tbfg_pcg_solve_queued_dev: %d iterations queued (expected 1..max_iter = %d)", iterations,
                  h->prm.max_iter);
        return TBFG_EINVAL;
    }
    if (h->fps) return solve_t<double, true>(h, psi, beta, q, nullptr, nullptr, stream, iterations);
    if (h->b16) return solve_t<float, false, bf16s>(h, psi, beta, q, nullptr, nullptr, stream, iterations);
    return h->mixed ? solve_t<float>(h, psi, beta, q, nullptr, nullptr, stream, iterations)
                    : solve_t<double>(h, psi, beta, q, nullptr, nullptr, stream, iterations);
}

int pcg_result(tbfg_pcg* h, int32_t* iters, double* res)
{
    if (!h) {
        set_error("tbfg_pcg_result: NULL handle");
        return TBFG_EINVAL;
    }
    TBFG_TRY(copy_sync(h->hst, h->st.p, sizeof(PcgState), hipMemcpyDeviceToHost));
    return finish_state(h, h->queued_fft, iters, res);
}

template <class T, class T0 = T>
static int precond_t(tbfg_pcg* h, const double* beta, const double* r, double* z, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    PcgState* st = h->st.as<PcgState>();
    Reducer red{h->partials.as<double>()};
    std::memset(h->hst, 0, sizeof(PcgState));
    TBFG_HIP(hipMemcpyAsync(st, h->hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    TBFG_TRY(setup_levels<T, T0>(h, beta, stream));
    const T0* beta0;
    if constexpr (sizeof(T) == 4)
        beta0 = h->beta[0].as<T0>();
    else
        beta0 = beta;
    const T0* r0;
    if constexpr (sizeof(T) == 4) {
        TBFG_TRY(launch_convert<double, T0>(h->n0, r, h->r32.as<T0>(), stream));
        r0 = h->r32.as<T0>();
    } else {
        r0 = r;
    }
    T* zt = nullptr;
    TBFG_TRY(vcycle<T, T0>(h, r0, &zt, beta0, nullptr, stream));
    // z with its arithmetic mean removed, as the reference's sweeps leave it
    TBFG_TRY(launch_pcg_dots<T>(h->n0, r, zt, nullptr, 1, st, red, stream));
    TBFG_TRY(launch_pcg_dir<T>(h->n0, zt, z, 1, st, stream));
    if (h->prm.mg_tol > 0.0) {  // the cycles taken, for tbfg_pcg_mg_stats (the loop synchronised already)
        TBFG_HIP(hipMemcpyAsync(h->hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        TBFG_HIP(hipStreamSynchronize(s));
    }
    return TBFG_OK;
}

int pcg_precondition_dev(tbfg_pcg* h, const double* beta, const double* r, double* z, void* stream)
{
    if (!h || !beta || !r || !z) {
        set_error("tbfg_pcg_precondition_dev: NULL argument");
        return TBFG_EINVAL;
    }
    if (h->fps) return fps_solve_dev(h->fps, r, z, stream);
    if (h->b16) return precond_t<float, bf16s>(h, beta, r, z, stream);
    return h->mixed ? precond_t<float>(h, beta, r, z, stream) : precond_t<double>(h, beta, r, z, stream);
}

}  // namespace tbfg
