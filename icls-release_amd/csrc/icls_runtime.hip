// icls_runtime.hip -- host side of libicls: the device context, parameter checks,
// constant tables, host<->device field transfer in the reference's Fortran layout,
// and the C-ABI entry points that stand in for the reference's subroutines.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "icls_internal.hpp"

namespace icls {

static Ctx* g_ctx = nullptr;
static char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------- profiling
struct Profiler {
    struct Rec { std::string name; hipEvent_t a, b; };
    std::vector<Rec> open;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, long>> acc;
    std::string only;   // when set, bracket only launches of this name
    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    void flush(hipStream_t s) {
        hipStreamSynchronize(s);
        for (auto& r : open) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, r.a, r.b);
            auto& x = acc[r.name];
            x.first += ms;
            x.second += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        open.clear();
    }
    ~Profiler() {
        for (auto& r : open) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
        for (auto e : pool) hipEventDestroy(e);
    }
};

void prof_begin(Ctx& c, const char* name) {
    Profiler* p = c.prof;
    if (!p->only.empty() && p->only != name) return;
    Profiler::Rec r{name, p->get(), p->get()};
    hipEventRecord(r.a, c.stream);
    p->open.push_back(r);
    if (p->open.size() > 4096) {
        // keep the event count bounded on long runs
        Profiler::Rec last = p->open.back();
        p->open.pop_back();
        p->flush(c.stream);
        p->open.push_back(last);
    }
}
void prof_end(Ctx& c, const char* name) {
    Profiler* p = c.prof;
    if (!p->only.empty() && p->only != name) return;
    for (auto it = p->open.rbegin(); it != p->open.rend(); ++it)
        if (it->name == name) { hipEventRecord(it->b, c.stream); return; }
}

// ---------------------------------------------------------------- tables
// pseudo_inv_A = (A^T A)^-1 A^T for the quadratic basis on the 27-point stencil
// (init.f90:1619-1650); Gauss-Jordan on the 10x10 normal matrix in long double,
// rounded once to double (the reference calls LAPACK dgetrf/dgetri).
void pseudo_inverse(double dx, double dy, double dz, double* out) {
    long double A[27][10], M[10][20];
    int cnt = 0;
    for (int dzn = -1; dzn <= 1; dzn++)
        for (int dyn = -1; dyn <= 1; dyn++)
            for (int dxn = -1; dxn <= 1; dxn++) {
                double a2 = dx * dxn, a3 = dy * dyn, a4 = dz * dzn;
                A[cnt][0] = 1.0L; A[cnt][1] = a2; A[cnt][2] = a3; A[cnt][3] = a4;
                A[cnt][4] = 0.5L * (long double)a2 * a2; A[cnt][5] = 0.5L * (long double)a3 * a3;
                A[cnt][6] = 0.5L * (long double)a4 * a4;
                A[cnt][7] = (long double)a2 * a3; A[cnt][8] = (long double)a2 * a4; A[cnt][9] = (long double)a3 * a4;
                cnt++;
            }
    for (int r = 0; r < 10; r++)
        for (int c = 0; c < 10; c++) {
            long double s = 0;
            for (int q = 0; q < 27; q++) s += A[q][r] * A[q][c];
            M[r][c] = s;
            M[r][10 + c] = (r == c) ? 1.0L : 0.0L;
        }
    for (int col = 0; col < 10; col++) {
        int piv = col;
        for (int r = col + 1; r < 10; r++) if (fabsl(M[r][col]) > fabsl(M[piv][col])) piv = r;
        if (piv != col) for (int c = 0; c < 20; c++) { long double t = M[col][c]; M[col][c] = M[piv][c]; M[piv][c] = t; }
        long double d = M[col][col];
        for (int c = 0; c < 20; c++) M[col][c] /= d;
        for (int r = 0; r < 10; r++) if (r != col) {
            long double f = M[r][col];
            if (f != 0) for (int c = 0; c < 20; c++) M[r][c] -= f * M[col][c];
        }
    }
    for (int r = 0; r < 10; r++)
        for (int q = 0; q < 27; q++) {
            long double s = 0;
            for (int c = 0; c < 10; c++) s += M[r][10 + c] * A[q][c];
            out[r + 10 * q] = (double)s;
        }
}

// eigenvalues of the periodic second difference in half-complex order
// (initsolver.f90:58-82)
static void eigenvalues(int n, double di, double pi, std::vector<double>& rt) {
    rt.assign(n + 1, 0.);
    for (int i = 3; i <= n; i += 2) {
        double s = sin((double)(i - 1) * pi / (2. * n));
        rt[i - 1] = -4. * di * di * (s * s);
        rt[i] = rt[i - 1];
    }
    rt[1] = 0.;
    rt[n] = -4. * di * di;
}

static bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// ---------------------------------------------------------------- context
static void free_ctx(Ctx* c) {
    if (!c) return;
    std::vector<void*> ptrs = {c->unew, c->vnew, c->wnew, c->dudt, c->dvdt, c->dwdt, c->pnew, c->pold, c->p,
                               c->p_jump, c->duold, c->dvold, c->dwold, c->p_x, c->p_y, c->p_z, c->p_xold,
                               c->p_yold, c->p_zold, c->miu, c->rho, c->curv, c->curv2, c->ucc, c->vcc, c->wcc,
                               c->u_l, c->v_l, c->w_l, c->dheav, c->ucc_l, c->vcc_l, c->wcc_l, c->normal_m,
                               c->ls_a, c->ls_b, c->ls_c, c->mask, c->thomas_d, c->dscal, c->dpart, c->pinvA,
                               c->xyrt, c->tw_x, c->tw_y, c->cos_tab, c->stash_b, c->dls, c->stash, c->scratch, c->pf, c->thomas_y,
                               c->thomas_seed, c->pinvA_f, c->tw_x_f, c->tw_y_f, c->pf0};
    if (c->frhs != c->p) ptrs.push_back(c->frhs);   // in FP64 the right-hand side is p itself
    ptrs.push_back(c->jmask);
    ptrs.push_back(c->jmask_old);
    ptrs.push_back(c->zplane);
    for (auto* p : ptrs) if (p) hipFree(p);
    for (auto* p : c->lvset) if (p) hipFree(p);
    if (c->hscal) hipHostFree(c->hscal);
    if (c->hvol) hipHostFree(c->hvol);
    if (c->dext) hipFree(c->dext);
    if (c->hext) hipHostFree(c->hext);
    if (c->dwin) hipFree(c->dwin);
    if (c->hwin) hipHostFree(c->hwin);
    if (c->ev_vol) hipEventDestroy(c->ev_vol);
    if (c->stream) hipStreamDestroy(c->stream);
    mlp_free(*c);
    delete c->prof;
    delete c;
}

// one padded 3D array of elements of esz bytes, zeroed
static int alloc_field(Ctx& c, double** p, size_t esz) {
    HIP_OK(hipMalloc((void**)p, esz * c.nfield));
    HIP_OK(hipMemset(*p, 0, esz * c.nfield));
    return ICLS_OK;
}

template <class T>
static int upload_table(const std::vector<double>& h, double** d) {
    std::vector<T> t(h.begin(), h.end());
    HIP_OK(hipMalloc((void**)d, sizeof(T) * t.size()));
    HIP_OK(hipMemcpy(*d, t.data(), sizeof(T) * t.size(), hipMemcpyHostToDevice));
    return ICLS_OK;
}

#define TRY(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

static int check_params(const icls_params* p) {
    if (!p) { set_error("null parameter block"); return ICLS_ERR_PARAM; }
    if (!pow2(p->itot) || !pow2(p->jtot) || p->itot < 8 || p->jtot < 8 || p->itot > 2048 || p->jtot > 2048) {
        set_error("itot and jtot must be powers of two in [8, 2048] (got %d, %d)", p->itot, p->jtot);
        return ICLS_ERR_PARAM;
    }
    if (p->ktot < 8) { set_error("ktot must be >= 8 (got %d)", p->ktot); return ICLS_ERR_PARAM; }
    if (p->lmax < 1 || p->lmax > ICLS_MAX_DROPS) { set_error("lmax out of range (%d)", p->lmax); return ICLS_ERR_PARAM; }
    if (!(p->lx > 0 && p->ly > 0 && p->lz > 0 && p->timestep > 0)) { set_error("domain and dt must be positive"); return ICLS_ERR_PARAM; }
    if (!(p->cfl >= 0. && p->cfl < 1e3)) { set_error("cfl must be in [0, 1e3) (got %g)", p->cfl); return ICLS_ERR_PARAM; }
    if (p->cfl > 0. && p->ioutchk < 1) { set_error("an adaptive dt (cfl > 0) needs ioutchk >= 1"); return ICLS_ERR_PARAM; }
    if (p->ls_adv_space != ICLS_ADV_HOUC5 && p->ls_adv_space != ICLS_ADV_WENO5) { set_error("unknown ls_adv_space"); return ICLS_ERR_PARAM; }
    if (p->mass_correction < 0 || p->mass_correction > 2) { set_error("mass_correction must be 0, 1 or 2"); return ICLS_ERR_PARAM; }
    if (p->reinit_step < 1 || p->inflate_step < 1) { set_error("reinit_step and inflate_step must be >= 1"); return ICLS_ERR_PARAM; }
    if (p->mass_trigger_exp < 0 || p->mass_trigger_exp > 15) {
        set_error("mass_trigger_exp must be in [0, 15] (got %d)", p->mass_trigger_exp);
        return ICLS_ERR_PARAM;
    }
    if (p->re_initialization && p->reinit_iter_max != 1) {
        set_error("reinit_iter_max other than 1 is not supported");
        return ICLS_ERR_UNSUPPORTED;
    }
    if (p->mode != ICLS_MODE_FP64 && p->mode != ICLS_MODE_MIXED && p->mode != ICLS_MODE_MIXED_LEARNED) {
        set_error("precision mode %d is not available", p->mode);
        return ICLS_ERR_UNSUPPORTED;
    }
    return ICLS_OK;
}

static int build_ctx(const icls_params* prm) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { set_error("no GPU visible"); return ICLS_ERR_NODEVICE; }
    Ctx* c = new Ctx();
    g_ctx = c;
    Ctx& C = *c;
    C.P = *prm;
    HIP_OK(hipGetDevice(&C.device));
    HIP_OK(hipDeviceGetAttribute(&C.ncu, hipDeviceAttributeMultiprocessorCount, C.device));
    HIP_OK(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
    TRY(noalloc_setup(C.device));   // the allocation-free chkdiv's partials on this GPU
    Grid& g = C.g;
    g.nx = prm->itot; g.ny = prm->jtot; g.nz = prm->ktot;
    g.sy = g.nx; g.sz = (long)g.nx * g.ny;
    // param.f90:37-38
    g.dxi = prm->itot / prm->lx; g.dyi = prm->jtot / prm->ly; g.dzi = prm->ktot / prm->lz;
    g.dx = 1. / g.dxi; g.dy = 1. / g.dyi; g.dz = 1. / g.dzi;
    auto p2 = [](double v) { int e; return std::frexp(v, &e) == 0.5; };
    g.pow2 = p2(g.dxi) && p2(g.dyi) && p2(g.dzi) && g.dx * g.dxi == 1. && g.dy * g.dyi == 1. && g.dz * g.dzi == 1.;
    C.lmax = prm->lmax;
    C.gfull = g;
    C.ncell_all = (long)g.nx * g.ny * g.nz;
    C.nfield = g.sz * (long)(g.nz + 6);
    C.NB2 = 6. * g.dz;
    C.rho0 = prm->rho1 < prm->rho2 ? prm->rho1 : prm->rho2;
    C.visc = 1. / prm->Re;
    C.pi = acos(-1.);
    C.dt = prm->timestep;
    C.fp32 = prm->mode != ICLS_MODE_FP64;
    C.learned = prm->mode == ICLS_MODE_MIXED_LEARNED;
    C.esz = C.fp32 ? sizeof(float) : sizeof(double);
    // field class: the mode's element type
    double** fields[] = {&C.unew, &C.vnew, &C.wnew, &C.dudt, &C.dvdt, &C.dwdt, &C.p_jump, &C.duold, &C.dvold,
                         &C.dwold, &C.p_x, &C.p_y, &C.p_z, &C.p_xold, &C.p_yold, &C.p_zold, &C.miu, &C.rho,
                         &C.curv, &C.curv2, &C.ucc, &C.vcc, &C.wcc, &C.u_l, &C.v_l, &C.w_l, &C.dheav, &C.ucc_l,
                         &C.vcc_l, &C.wcc_l, &C.normal_m, &C.ls_a, &C.ls_b, &C.ls_c};
    for (auto** f : fields) TRY(alloc_field(C, f, C.esz));
    C.lvset.assign(C.lmax, nullptr);
    for (int l = 0; l < C.lmax; l++) TRY(alloc_field(C, &C.lvset[l], C.esz));
    // pressure class: double in every mode
    double** pfields[] = {&C.pnew, &C.pold, &C.p, &C.scratch};
    for (auto** f : pfields) TRY(alloc_field(C, f, sizeof(double)));
    C.thomas_seg = thomas_seg_len(g.nz);
    if (C.thomas_seg) {
        const long nseed = (long)g.nx * g.ny * (g.nz / C.thomas_seg - 1);
        HIP_OK(hipMalloc((void**)&C.thomas_seed, sizeof(double) * (nseed > 0 ? nseed : 1)));
    } else {
        TRY(alloc_field(C, &C.thomas_d, sizeof(double)));
        if (C.fp32) TRY(alloc_field(C, &C.thomas_y, sizeof(double)));
    }
    if (C.fp32) {
        TRY(alloc_field(C, &C.frhs, sizeof(float)));
        TRY(alloc_field(C, &C.pf, sizeof(float)));
        TRY(alloc_field(C, &C.pf0, sizeof(float)));
    } else {
        C.frhs = C.p;
    }
    HIP_OK(hipMalloc((void**)&C.mask, C.nfield));
    HIP_OK(hipMemset(C.mask, 0, C.nfield));
    // face-jump masks start with every bit set: the (zero) arrays are read as they are
    HIP_OK(hipMalloc((void**)&C.jmask, C.nfield));
    HIP_OK(hipMemset(C.jmask, 7, C.nfield));
    HIP_OK(hipMalloc((void**)&C.jmask_old, C.nfield));
    HIP_OK(hipMemset(C.jmask_old, 7, C.nfield));
    HIP_OK(hipMalloc((void**)&C.zplane, sizeof(double) * g.nx * g.ny));
    HIP_OK(hipMemset(C.zplane, 0, sizeof(double) * g.nx * g.ny));
    HIP_OK(hipMalloc((void**)&C.stash, sizeof(double) * 2 * g.sz));
    // volumes, initial volumes and scratch sums per level set, then the batched
    // re-init residual slot of each level set and two saved scalars
    const int nscal = DS_COUNT_FIXED + 3 * C.lmax + 16 + C.lmax + 2;
    C.res_slot = DS_REINIT_RES;
    C.res_b0 = DS_COUNT_FIXED + 3 * C.lmax + 16;
    C.save0 = C.res_b0 + C.lmax;
    HIP_OK(hipMalloc((void**)&C.dscal, sizeof(double) * nscal));
    HIP_OK(hipMemset(C.dscal, 0, sizeof(double) * nscal));
    HIP_OK(hipHostMalloc((void**)&C.hscal, sizeof(double) * nscal));
    memset(C.hscal, 0, sizeof(double) * nscal);
    C.nscal = nscal;
    HIP_OK(hipHostMalloc((void**)&C.hvol, sizeof(double) * 2 * C.lmax));
    memset(C.hvol, 0, sizeof(double) * 2 * C.lmax);
    HIP_OK(hipEventCreateWithFlags(&C.ev_vol, hipEventDisableTiming));
    if (C.lmax > 1 && !prm->ls_full) {
        HIP_OK(hipMalloc((void**)&C.dext, sizeof(int) * 6 * C.lmax));
        HIP_OK(hipHostMalloc((void**)&C.hext, sizeof(int) * 6 * C.lmax));
        HIP_OK(hipMalloc((void**)&C.dwin, sizeof(int) * 6 * C.lmax));
        HIP_OK(hipHostMalloc((void**)&C.hwin, sizeof(int) * 6 * C.lmax));
        HIP_OK(hipMalloc((void**)&C.stash_b, (size_t)C.esz * 2 * g.sz * WB_MAX));
    }
    // reduction partials: enough for lmax per-block sums of any sweep
    // (the chkdt sweep covers 1..i1 x 1..j1 x 1..k1, one row and column more than the others)
    long nb = (long)((g.nx + 2 + 63) / 64 + 1) * ((g.ny + 2 + 3) / 4 + 1) * ((g.nz + 2 + 7) / 8 + 1);
    C.npart_cap = nb * (C.lmax > 3 ? C.lmax : 3) + 4096;
    HIP_OK(hipMalloc((void**)&C.dpart, sizeof(double) * C.npart_cap));
    HIP_OK(hipMalloc((void**)&C.dls, sizeof(double*) * C.lmax));
    HIP_OK(hipMemcpy(C.dls, C.lvset.data(), sizeof(double*) * C.lmax, hipMemcpyHostToDevice));
    // constant tables
    C.host_pinvA.assign(270, 0.);
    pseudo_inverse(g.dx, g.dy, g.dz, C.host_pinvA.data());
    TRY(upload_table<double>(C.host_pinvA, &C.pinvA));
    if (C.fp32) TRY(upload_table<float>(C.host_pinvA, &C.pinvA_f));
    std::vector<double> xrt, yrt;
    eigenvalues(g.nx, g.dxi, C.pi, xrt);
    eigenvalues(g.ny, g.dyi, C.pi, yrt);
    std::vector<double> xy((size_t)g.nx * g.ny);
    for (int j = 1; j <= g.ny; j++)
        for (int i = 1; i <= g.nx; i++) xy[(size_t)(i - 1) + (size_t)g.nx * (j - 1)] = xrt[i] + yrt[j];
    HIP_OK(hipMalloc((void**)&C.xyrt, sizeof(double) * xy.size()));
    HIP_OK(hipMemcpy(C.xyrt, xy.data(), sizeof(double) * xy.size(), hipMemcpyHostToDevice));
    TRY(launch_thomas_seed(C));
    for (int dir = 0; dir < 2; dir++) {
        int n = dir == 0 ? g.nx : g.ny;
        std::vector<double> tw((size_t)n);   // n/2 complex exp(+2 pi i t/n)
        for (int t = 0; t < n / 2; t++) {
            double ang = 2. * C.pi * (double)t / (double)n;
            tw[2 * t] = cos(ang);
            tw[2 * t + 1] = sin(ang);
        }
        TRY(upload_table<double>(tw, dir == 0 ? &C.tw_x : &C.tw_y));
        if (C.fp32) TRY(upload_table<float>(tw, dir == 0 ? &C.tw_x_f : &C.tw_y_f));
    }
    TRY(mlp_alloc(C));
    HIP_OK(hipDeviceSynchronize());
    return ICLS_OK;
}

int zchunk_for(Ctx& c, const void* fn, int threads, int tiles_xy, int mult, int halo) {
    auto it = c.zchunk.find(fn);
    if (it != c.zchunk.end()) return it->second;
    const int nz = c.g.nz;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, 0) != hipSuccess || occ < 1) occ = 1;
    const long slots = (long)occ * (c.ncu > 0 ? c.ncu : 1);
    if (mult < 1 || nz % mult != 0) mult = 1;
    int best = nz;
    long best_cost = -1;
    for (int ntz = 1; ntz <= nz; ntz++) {
        int ch = (nz + ntz - 1) / ntz;
        ch = (ch + mult - 1) / mult * mult;
        if ((nz + ch - 1) / ch != ntz) continue;           // each chunk count once
        if (ch < (nz < 16 ? nz : 16)) break;               // chunks of 16 layers at least
        const long blocks = (long)tiles_xy * ntz;
        const long cost = (blocks + slots - 1) / slots * (ch + halo);
        if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = ch; }
    }
    c.zchunk[fn] = best;
    return best;
}

static Ctx* need_ctx() {
    if (!g_ctx) set_error("icls_init has not been called");
    return g_ctx;
}

#define CTX_OR_FAIL()            \
    Ctx* cp = need_ctx();        \
    if (!cp) return ICLS_ERR_STATE; \
    Ctx& c = *cp

// ---------------------------------------------------------------- field layout conversion
static double* field_ptr(Ctx& c, int field, int l) {
    switch (field) {
    case ICLS_F_UNEW: return c.unew;  case ICLS_F_VNEW: return c.vnew;  case ICLS_F_WNEW: return c.wnew;
    case ICLS_F_DUDT: return c.dudt;  case ICLS_F_DVDT: return c.dvdt;  case ICLS_F_DWDT: return c.dwdt;
    case ICLS_F_PNEW: return c.pnew;  case ICLS_F_POLD: return c.pold;  case ICLS_F_P: return c.p;
    case ICLS_F_PJUMP: return c.p_jump;
    case ICLS_F_DUOLD: return c.duold; case ICLS_F_DVOLD: return c.dvold; case ICLS_F_DWOLD: return c.dwold;
    case ICLS_F_PX: return c.p_x; case ICLS_F_PY: return c.p_y; case ICLS_F_PZ: return c.p_z;
    case ICLS_F_PXOLD: return c.p_xold; case ICLS_F_PYOLD: return c.p_yold; case ICLS_F_PZOLD: return c.p_zold;
    case ICLS_F_MIU: return c.miu; case ICLS_F_RHO: return c.rho; case ICLS_F_CURV: return c.curv2;
    case ICLS_F_UCC: return c.ucc; case ICLS_F_VCC: return c.vcc; case ICLS_F_WCC: return c.wcc;
    case ICLS_F_LVSET: return (l >= 0 && l < c.lmax) ? c.lvset[l] : nullptr;
    case ICLS_F_NORMAL_M: return c.normal_m;
    default: return nullptr;
    }
}

static int shape_of(int field) {
    switch (field) {
    case ICLS_F_DUOLD: case ICLS_F_DVOLD: case ICLS_F_DWOLD: case ICLS_F_UCC: case ICLS_F_VCC: case ICLS_F_WCC:
    case ICLS_F_RHO_U: case ICLS_F_RHO_V: case ICLS_F_RHO_W: case ICLS_F_NORMAL_M: return ICLS_SHAPE_I;
    case ICLS_F_LVSET: return ICLS_SHAPE_LS;
    default: return ICLS_SHAPE_C;
    }
}

// pressure-class fields are double in every mode; the others hold the mode's type
static bool field_is_fp32(const Ctx& c, int field) {
    if (!c.fp32) return false;
    return !(field == ICLS_F_PNEW || field == ICLS_F_POLD || field == ICLS_F_P || field == ICLS_F_JIEMIAN);
}

// whole padded device array -> host doubles
static int fetch_array(Ctx& c, const double* d, bool fp32, std::vector<double>& out) {
    out.resize((size_t)c.nfield);
    if (fp32) {
        std::vector<float> f((size_t)c.nfield);
        HIP_OK(hipMemcpyAsync(f.data(), d, sizeof(float) * c.nfield, hipMemcpyDeviceToHost, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
        for (size_t q = 0; q < f.size(); q++) out[q] = (double)f[q];
    } else {
        HIP_OK(hipMemcpyAsync(out.data(), d, sizeof(double) * c.nfield, hipMemcpyDeviceToHost, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    }
    return ICLS_OK;
}

// host doubles -> whole padded device array (rounded to float when fp32)
static int store_array(Ctx& c, double* d, bool fp32, const std::vector<double>& in) {
    if (fp32) {
        std::vector<float> f(in.begin(), in.end());
        HIP_OK(hipMemcpyAsync(d, f.data(), sizeof(float) * c.nfield, hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    } else {
        HIP_OK(hipMemcpyAsync(d, in.data(), sizeof(double) * c.nfield, hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    }
    return ICLS_OK;
}

// host Fortran extents and the offsets mapping host (ih,jh,kh) to (Fortran i,j,k)
static void host_dims(const Ctx& c, int shape, int& hx, int& hy, int& hz, int& lb) {
    const Grid& g = c.g;
    if (shape == ICLS_SHAPE_C) { hx = g.nx + 2; hy = g.ny + 2; hz = g.nz + 2; lb = 0; }
    else if (shape == ICLS_SHAPE_LS) { hx = g.nx + 6; hy = g.ny + 6; hz = g.nz + 6; lb = -2; }
    else { hx = g.nx; hy = g.ny; hz = g.nz; lb = 1; }
}

}  // namespace icls

using namespace icls;

extern "C" {

const char* icls_last_error(void) { return g_err; }
int32_t icls_version(void) { return 100; }

int32_t icls_init(const icls_params* p) {
    int rc = check_params(p);
    if (rc) return rc;
    if (g_ctx) { free_ctx(g_ctx); g_ctx = nullptr; }
    rc = build_ctx(p);
    if (rc) { free_ctx(g_ctx); g_ctx = nullptr; }
    return rc;
}

int32_t icls_finalize(void) {
    if (g_ctx) { hipStreamSynchronize(g_ctx->stream); free_ctx(g_ctx); g_ctx = nullptr; }
    return ICLS_OK;
}

int32_t icls_field_shape(int32_t field, int32_t* shape, int64_t* count) {
    CTX_OR_FAIL();
    if (field < 0 || field >= ICLS_F_COUNT) { set_error("bad field id %d", field); return ICLS_ERR_PARAM; }
    int sh = shape_of(field), hx, hy, hz, lb;
    host_dims(c, sh, hx, hy, hz, lb);
    if (shape) *shape = sh;
    if (count) *count = (int64_t)hx * hy * hz;
    return ICLS_OK;
}

int32_t icls_sync(void) {
    CTX_OR_FAIL();
    HIP_OK(hipStreamSynchronize(c.stream));
    return ICLS_OK;
}

// the pressures of a lazy rotation are made stored values before they are accessed
static int prepare_field(Ctx& c, int field) {
    if (field == ICLS_F_PNEW || field == ICLS_F_POLD) return launch_pressure_materialize(c);
    if (field == ICLS_F_PX || field == ICLS_F_PY || field == ICLS_F_PZ) return launch_jump_densify(c, false);
    if (field == ICLS_F_PXOLD || field == ICLS_F_PYOLD || field == ICLS_F_PZOLD) return launch_jump_densify(c, true);
    if (field == ICLS_F_P) return launch_pressure_release_p(c);
    return ICLS_OK;
}

static int init_windows(Ctx& c);

int32_t icls_upload(int32_t field, int32_t l, const double* host) {
    CTX_OR_FAIL();
    c.vbulk_fresh = false;
    TRY(prepare_field(c, field));
    double* d = field_ptr(c, field, l);
    if (!d || !host) { set_error("field %d (l=%d) cannot be uploaded", field, l); return ICLS_ERR_PARAM; }
    int sh = shape_of(field), hx, hy, hz, lb;
    host_dims(c, sh, hx, hy, hz, lb);
    const Grid& g = c.g;
    const bool f32 = field_is_fp32(c, field);
    std::vector<double> stage;
    TRY(fetch_array(c, d, f32, stage));
    for (int kh = 0; kh < hz; kh++) {
        int k = kh + lb;   // Fortran k
        for (int j0 = 0; j0 < g.ny; j0++)
            for (int i0 = 0; i0 < g.nx; i0++) {
                int ih = i0 + 1 - lb, jh = j0 + 1 - lb;
                stage[(size_t)gidx(g, i0, j0, k)] = host[(size_t)ih + (size_t)hx * ((size_t)jh + (size_t)hy * kh)];
            }
    }
    TRY(store_array(c, d, f32, stage));
    // in the mixed modes the solve reads its right-hand side from a float array
    if (field == ICLS_F_P && c.frhs != c.p) TRY(store_array(c, c.frhs, true, stage));
    // a new level set: its window is found again from the whole grid
    if (field == ICLS_F_LVSET && !c.win.empty()) TRY(init_windows(c));
    return ICLS_OK;
}

int32_t icls_download(int32_t field, int32_t l, double* host) {
    CTX_OR_FAIL();
    if (!host) { set_error("null host buffer"); return ICLS_ERR_PARAM; }
    TRY(prepare_field(c, field));
    double* d = field_ptr(c, field, l);
    if (field == ICLS_F_JIEMIAN) { TRY(launch_jiemian(c, c.scratch)); d = c.scratch; }
    if (field == ICLS_F_UCC || field == ICLS_F_VCC || field == ICLS_F_WCC) TRY(launch_avg_vel(c));
    if (field == ICLS_F_RHO_U || field == ICLS_F_RHO_V || field == ICLS_F_RHO_W) d = c.rho;
    if (!d) { set_error("field %d (l=%d) cannot be downloaded", field, l); return ICLS_ERR_PARAM; }
    int sh = shape_of(field), hx, hy, hz, lb;
    host_dims(c, sh, hx, hy, hz, lb);
    const Grid& g = c.g;
    std::vector<double> stage;
    TRY(fetch_array(c, d, field_is_fp32(c, field), stage));
    const bool face = field == ICLS_F_RHO_U || field == ICLS_F_RHO_V || field == ICLS_F_RHO_W;
    const bool req = c.P.rho1 == c.P.rho2;
    // gfm4p zeroes p_jump and the face jumps and fills the interior only (gfm.f90:71-79,
    // 219-233); nothing exchanges their halos, which stay zero in the reference
    const bool jump = field == ICLS_F_PJUMP || field == ICLS_F_PX || field == ICLS_F_PY || field == ICLS_F_PZ ||
                      field == ICLS_F_PXOLD || field == ICLS_F_PYOLD || field == ICLS_F_PZOLD;
    for (int kh = 0; kh < hz; kh++) {
        int k = kh + lb;
        for (int jh = 0; jh < hy; jh++) {
            int j0 = wrap(jh + lb - 1, g.ny);
            const bool jhalo = jh + lb - 1 < 0 || jh + lb - 1 >= g.ny;
            for (int ih = 0; ih < hx; ih++) {
                int i0 = wrap(ih + lb - 1, g.nx);
                double v = stage[(size_t)gidx(g, i0, j0, k)];
                if (jump && (jhalo || ih + lb - 1 < 0 || ih + lb - 1 >= g.nx)) v = 0.;
                if (face) {
                    // arithmetic face means (misc.f90:151-172)
                    if (req) v = c.P.rho1;
                    else if (field == ICLS_F_RHO_U) v = (v + stage[(size_t)gidx(g, wrap(i0 + 1, g.nx), j0, k)]) / 2.;
                    else if (field == ICLS_F_RHO_V) v = (v + stage[(size_t)gidx(g, i0, wrap(j0 + 1, g.ny), k)]) / 2.;
                    else v = (v + stage[(size_t)gidx(g, i0, j0, k + 1)]) / 2.;
                }
                host[(size_t)ih + (size_t)hx * ((size_t)jh + (size_t)hy * kh)] = v;
            }
        }
    }
    return ICLS_OK;
}

static int pull_scalars(Ctx& c) {
    HIP_OK(hipMemcpyAsync(c.hscal, c.dscal, sizeof(double) * c.nscal, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return ICLS_OK;
}

int32_t icls_get_scalar(int32_t which, double* value) {
    CTX_OR_FAIL();
    if (!value) return ICLS_ERR_PARAM;
    if (which == ICLS_S_TIME) { *value = c.time; return ICLS_OK; }
    if (which == ICLS_S_DT) { *value = c.dt; return ICLS_OK; }
    TRY(pull_scalars(c));
    switch (which) {
    case ICLS_S_RHO_AVR: *value = c.hscal[DS_RHO_AVR]; break;
    case ICLS_S_V_BULK: *value = c.hscal[DS_V_BULK] / (c.g.nx * c.g.ny * c.g.nz); break;
    case ICLS_S_MAXCHK: *value = c.hscal[DS_MAXCHK]; break;
    case ICLS_S_REINIT_RES: *value = c.hscal[DS_REINIT_RES]; break;
    case ICLS_S_NORM_OLD: *value = c.hscal[DS_NORM_OLD]; break;
    default: set_error("bad scalar id %d", which); return ICLS_ERR_PARAM;
    }
    return ICLS_OK;
}

int32_t icls_set_scalar(int32_t which, double value) {
    CTX_OR_FAIL();
    int slot = -1;
    switch (which) {
    case ICLS_S_TIME: c.time = value; return ICLS_OK;
    case ICLS_S_DT: c.dt = value; return ICLS_OK;
    case ICLS_S_RHO_AVR: slot = DS_RHO_AVR; break;
    case ICLS_S_NORM_OLD: slot = DS_NORM_OLD; break;
    default: set_error("scalar %d is not settable", which); return ICLS_ERR_PARAM;
    }
    HIP_OK(hipMemcpyAsync(c.dscal + slot, &value, sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return ICLS_OK;
}

int32_t icls_get_volume(int32_t l, double* volume, double* init_volume) {
    CTX_OR_FAIL();
    if (l < 0 || l >= c.lmax) { set_error("bad level-set index %d", l); return ICLS_ERR_PARAM; }
    TRY(pull_scalars(c));
    if (volume) *volume = c.hscal[DS_COUNT_FIXED + l];
    if (init_volume) *init_volume = c.hscal[DS_COUNT_FIXED + c.lmax + l];
    return ICLS_OK;
}

static int publish_volumes(Ctx& c);
static int train_impl(Ctx& c, icls_mlp_train_opts* o);

int32_t icls_set_volume(int32_t l, double volume, double init_volume) {
    CTX_OR_FAIL();
    if (l < 0 || l >= c.lmax) { set_error("bad level-set index %d", l); return ICLS_ERR_PARAM; }
    double v[1] = {volume};
    HIP_OK(hipMemcpyAsync(c.dscal + DS_COUNT_FIXED + l, v, sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    v[0] = init_volume;
    HIP_OK(hipMemcpyAsync(c.dscal + DS_COUNT_FIXED + c.lmax + l, v, sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return publish_volumes(c);
}

// The mass-correction trigger needs volume(l) from the previous update_materials.
// The volumes are copied to pinned host memory right after they are formed and the
// host waits for that copy only when it reaches the trigger, by which time the device
// is still working through the queued Navier-Stokes half of the previous step.
static int publish_volumes(Ctx& c) {
    HIP_OK(hipMemcpyAsync(c.hvol, c.dscal + DS_COUNT_FIXED, sizeof(double) * 2 * c.lmax, hipMemcpyDeviceToHost,
                          c.stream));
    HIP_OK(hipEventRecord(c.ev_vol, c.stream));
    c.vol_published = true;
    return ICLS_OK;
}

// ls_mass_correction trigger (levelset.f90:386-395), evaluated on the host copy of
// the same doubles the device flag kernel uses
static int mass_correction_due(Ctx& c, int l, int istep, bool& due) {
    due = false;
    if (istep % c.P.inflate_step != 0) return ICLS_OK;
    if (!c.vol_published) TRY(publish_volumes(c));
    HIP_OK(hipEventSynchronize(c.ev_vol));
    double V = c.hvol[l], V0 = c.hvol[c.lmax + l];
    double massloss = (V - V0) / V0;
    due = fabs(massloss) > mass_trigger(c.P);
    return ICLS_OK;
}

// ---------------------------------------------------------------- level set
// ---------------------------------------------------------------- level-set windows
// With several level sets each one is stepped inside a window around its drop
// (LsWindow): the launches see a grid of the window's size whose strides are the whole
// grid's and whose arrays start at the window's first cell.  x and y wrap inside the
// window, so a window is a small periodic domain; its edges lie in the clamped far
// field (phi = NarrowBand_2 after the tube), where every stencil is flat, and the
// stepping inside equals the whole-grid stepping wherever phi is below NarrowBand_2.
// The arrays keep their whole-grid size: cells outside a window keep the values they
// had when they left it (far field, phi >= NarrowBand_2, Heaviside 1).
constexpr int WIN_REFRESH = 4;     // ls_main passes between window updates
constexpr int WIN_REACH = 7;       // stencil reach of one step beyond the band: tube scan,
                                   // RK stages, re-init
constexpr int WIN_DRIFT = 2;       // band motion allowed between updates (cells)

struct WindowScope {
    Ctx& c;
    bool on;
    int l;
    Grid g{};
    double* saved[21];
    double* ls = nullptr;
    uint8_t* mask = nullptr;
    double* stash = nullptr;
    double** fld(int i) {
        double** f[21] = {&c.unew, &c.vnew, &c.wnew, &c.curv, &c.curv2, &c.ucc_l, &c.vcc_l, &c.wcc_l, &c.u_l,
                          &c.v_l, &c.w_l, &c.dheav, &c.normal_m, &c.ls_a, &c.ls_b, &c.ls_c, &c.p_jump, &c.p_x,
                          &c.p_y, &c.p_z, &c.ucc};
        return f[i];
    }
    static double* shift(double* p, long off, size_t esz) { return (double*)((char*)p + off * (long)esz); }
    WindowScope(Ctx& c_, int l_) : c(c_), on(!c_.win.empty()), l(l_) {
        if (!on) return;
        const LsWindow& w = c.win[l];
        g = c.g;
        const long poff = w.ox + c.g.sy * (long)w.oy, off = poff + c.g.sz * (long)w.oz;
        c.g.nx = w.wx; c.g.ny = w.wy; c.g.nz = w.wz;
        for (int i = 0; i < 21; i++) { saved[i] = *fld(i); *fld(i) = shift(saved[i], off, c.esz); }
        ls = c.lvset[l]; c.lvset[l] = shift(ls, off, c.esz);
        mask = c.mask; c.mask = mask + off;
        stash = c.stash; c.stash = shift(stash, poff, c.esz);
    }
    ~WindowScope() {
        if (!on) return;
        c.g = g;
        for (int i = 0; i < 21; i++) *fld(i) = saved[i];
        c.lvset[l] = ls;
        c.mask = mask;
        c.stash = stash;
    }
};

// smallest power of two >= v
static int pow2_at_least(int v) { int p = 1; while (p < v) p <<= 1; return p; }

// One periodic axis: cover global cells lo .. hi (lo may be < 0 or hi >= n when the
// drop crosses the boundary) with a range that does not cross it, else with the whole
// axis.  x ranges are powers of two (the z-marching kernels wrap x with a mask, and
// their tiles are 64 cells wide); y ranges are multiples of 8, which on drops of radius
// 3 .. 20 cells cut the cells of the windows by a quarter against powers of two.
static void window_axis_periodic(int lo, int hi, int n, bool pow2_range, int& o, int& w) {
    const int need = hi - lo + 1;
    w = pow2_range ? pow2_at_least(need) : (need + 7) / 8 * 8;
    if (lo < 0 || hi > n - 1 || w >= n) { o = 0; w = n; return; }
    o = lo - (w - need) / 2;
    if (o < 0) o = 0;
    if (o > n - w) o = n - w;
}

// z (walls): layers lo .. hi clamped to 1 .. nz, wz a multiple of 4 where nz allows
static void window_axis_z(int lo, int hi, int n, int& o, int& w) {
    if (lo < 1) lo = 1;
    if (hi > n) hi = n;
    w = hi - lo + 1;
    if (n % 4 == 0) {
        w = (w + 3) / 4 * 4;
        if (lo - 1 + w > n) lo = n - w + 1;
    }
    if (w >= n) { o = 0; w = n; return; }
    o = lo - 1;
}

// Windows from the current level sets: the bounding box of phi < NarrowBand_2 inside
// each window, widened by WIN_REACH + WIN_DRIFT cells.  Fails when a drop's band came
// closer than WIN_REACH cells to an edge of its window that is not a wall or the whole
// periodic axis.
static int update_windows(Ctx& c) {
    const int L = c.lmax;
    HIP_OK(hipMemsetAsync(c.dext, 0x7f, sizeof(int) * 6 * L, c.stream));
    for (int l = 0; l < L; l++) {
        WindowScope ws(c, l);
        TRY(launch_ls_extent(c, c.lvset[l], c.dext + 6 * l));
    }
    HIP_OK(hipMemcpyAsync(c.hext, c.dext, sizeof(int) * 6 * L, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    const Grid& G = c.gfull;
    const int m = WIN_REACH + WIN_DRIFT;
    for (int l = 0; l < L; l++) {
        LsWindow& w = c.win[l];
        const int* e = c.hext + 6 * l;
        if (e[0] == 0x7f7f7f7f) continue;          // no cell below the band limit: keep it
        const int lo[3] = {e[0], e[1], e[2]}, hi[3] = {-e[3], -e[4], -e[5]};
        const int ext[3] = {w.wx, w.wy, w.wz}, full[3] = {G.nx, G.ny, G.nz};
        for (int a = 0; a < 3; a++) {
            const bool whole = ext[a] == full[a];
            const int first = a == 2 ? 1 : 0, last = a == 2 ? ext[a] : ext[a] - 1;
            const bool wall_lo = a == 2 && w.oz == 0, wall_hi = a == 2 && w.oz + w.wz == G.nz;
            if (!whole && ((!wall_lo && lo[a] - first < WIN_REACH) || (!wall_hi && last - hi[a] < WIN_REACH))) {
                set_error("level set %d came within %d cells of its window edge; the band moves faster than "
                          "%d cells per %d steps", l, WIN_REACH, WIN_DRIFT, WIN_REFRESH);
                return ICLS_ERR_STATE;
            }
        }
        LsWindow n;
        if (w.wx == G.nx && (lo[0] < m || hi[0] > G.nx - 1 - m)) { n.ox = 0; n.wx = G.nx; }
        else window_axis_periodic(w.ox + lo[0] - m, w.ox + hi[0] + m, G.nx, true, n.ox, n.wx);
        if (w.wy == G.ny && (lo[1] < m || hi[1] > G.ny - 1 - m)) { n.oy = 0; n.wy = G.ny; }
        else window_axis_periodic(w.oy + lo[1] - m, w.oy + hi[1] + m, G.ny, false, n.oy, n.wy);
        window_axis_z(w.oz + lo[2] - m, w.oz + hi[2] + m, G.nz, n.oz, n.wz);
        w = n;
    }
    // colours: windows of one colour share no cell they read or write -- x and y inside
    // the window, z from 2 layers below to 3 above it (the ghost layers)
    auto meets = [](const LsWindow& a, const LsWindow& b) {
        return a.ox < b.ox + b.wx && b.ox < a.ox + a.wx && a.oy < b.oy + b.wy && b.oy < a.oy + a.wy &&
               a.oz - 2 <= b.oz + b.wz + 3 && b.oz - 2 <= a.oz + a.wz + 3;
    };
    c.colour.assign(L, 0);
    c.ncolour = 0;
    std::vector<char> used;
    for (int l = 0; l < L; l++) {
        used.assign(c.ncolour + 1, 0);
        for (int m = 0; m < l; m++)
            if (meets(c.win[l], c.win[m])) used[c.colour[m]] = 1;
        int k = 0;
        while (used[k]) k++;
        c.colour[l] = k;
        if (k + 1 > c.ncolour) c.ncolour = k + 1;
    }
    // the table k_materials reads (the copy completes before the next update writes hwin:
    // the update synchronises the stream first)
    for (int l = 0; l < L; l++) {
        const LsWindow& w = c.win[l];
        int* h = c.hwin + 6 * l;
        h[0] = w.ox; h[1] = w.oy; h[2] = w.oz; h[3] = w.wx; h[4] = w.wy; h[5] = w.wz;
    }
    HIP_OK(hipMemcpyAsync(c.dwin, c.hwin, sizeof(int) * 6 * L, hipMemcpyHostToDevice, c.stream));
    return ICLS_OK;
}

// windows on (several level sets, not ls_full): start from the whole grid
static int init_windows(Ctx& c) {
    c.win.clear();
    if (!c.dext) return ICLS_OK;
    LsWindow all;
    all.wx = c.gfull.nx; all.wy = c.gfull.ny; all.wz = c.gfull.nz;
    c.win.assign(c.lmax, all);
    c.win_age = 0;
    return update_windows(c);
}

static int ls_main_impl(Ctx& c, int istep) {
    if (!c.win.empty() && ++c.win_age >= WIN_REFRESH) {
        TRY(update_windows(c));
        c.win_age = 0;
    }
    const int sch = c.P.ls_adv_space;
    // tube_construction (levelset.f90:87-197) and ls_advection, SSP-RK3
    // (levelset.f90:235-243 / 268-276), cell-centre velocity averaged from
    // unew/vnew/wnew inside the stencil (misc.f90:595-597); every level-set kernel
    // folds the boundsls that follows it (bound.f90:404-411) into its own launch
    auto advect = [&](int l) -> int {
        double* ph = c.lvset[l];
        TRY(launch_tube(c, ph, c.ls_a));
        TRY(launch_adv_stage(c, sch, 1, c.ls_a, c.ls_a, c.ls_b, c.unew, c.vnew, c.wnew, 1, nullptr));
        TRY(launch_adv_stage(c, sch, 2, c.ls_a, c.ls_b, c.ls_c, c.unew, c.vnew, c.wnew, 1, nullptr));
        return launch_adv_stage(c, sch, 3, c.ls_a, c.ls_c, ph, c.unew, c.vnew, c.wnew, 1, nullptr);
    };
    // ls_mass_correction (levelset.f90:377-582)
    auto correct = [&](int l) -> int {
        if (!c.P.mass_correction) return ICLS_OK;
        double* ph = c.lvset[l];
        bool due = false;
        TRY(mass_correction_due(c, l, istep, due));
        // maxchk_glb = 0 and the device flag; when the correction is not due the call
        // only resets maxchk, which the last level set's call leaves as the value
        if (due || l == c.lmax - 1) TRY(launch_inflation_prepare(c, l, istep));
        if (!due) return ICLS_OK;
        const double* flag = c.dscal + DS_INFL_FLAG;
        TRY(launch_curvature(c, ph, c.curv2, c.normal_m, flag));
        TRY(launch_inflation_velocity(c, l));
        TRY(launch_adv_stage(c, ICLS_ADV_HOUC5, 1, ph, ph, c.ls_a, c.ucc_l, c.vcc_l, c.wcc_l, 0, flag));
        TRY(launch_adv_stage(c, ICLS_ADV_HOUC5, 2, ph, c.ls_a, c.ls_b, c.ucc_l, c.vcc_l, c.wcc_l, 0, flag));
        return launch_adv_stage(c, ICLS_ADV_HOUC5, 3, ph, c.ls_b, ph, c.ucc_l, c.vcc_l, c.wcc_l, 0, flag);
    };
    // ls_reinit_main, WENO5 + RK3 (levelset.f90:592-882)
    const bool reinit = c.P.re_initialization && istep % c.P.reinit_step == 0;
    auto reinitialise = [&](int l) -> int {
        double* ph = c.lvset[l];
        TRY(launch_reinit_stage(c, 1, ph, ph, c.ls_a));
        TRY(launch_reinit_stage(c, 2, ph, c.ls_a, c.ls_b));
        return launch_reinit_stage(c, 3, ph, c.ls_b, ph);    // in place: lvset = phi_new
    };
    if (c.win.empty()) {
        for (int l = 0; l < c.lmax; l++) {
            TRY(advect(l));
            TRY(correct(l));
            if (reinit) TRY(reinitialise(l));
        }
    } else {
        // Colour by colour, all windows of a colour in one launch per kernel: advection,
        // then the colour's mass corrections one at a time, then (after every colour)
        // the re-initialisation.  A correction advects inside the band mask of its own
        // level set's tube, as the reference's loop does (levelset.f90:47-68); the mask
        // array is shared, and a later colour's tube rewrites it where windows overlap,
        // so each colour is corrected before the next colour's tubes run.
        std::vector<std::vector<int>> by_colour(c.ncolour);
        for (int l = 0; l < c.lmax; l++) by_colour[c.colour[l]].push_back(l);
        // the reference ends the loop with maxchk and the flag of the last level set's
        // correction: they are kept aside when a later colour's trigger overwrites them
        bool saved = false, later = false;
        const int slots[2] = {DS_MAXCHK, DS_INFL_FLAG};
        for (const auto& ls : by_colour) {
            if (ls.empty()) continue;
            TRY(launch_advect_batch(c, ls));
            if (!c.P.mass_correction) continue;
            for (int l : ls) {
                {
                    WindowScope ws(c, l);
                    TRY(correct(l));
                }
                if (l == c.lmax - 1) {
                    for (int q = 0; q < 2; q++)
                        HIP_OK(hipMemcpyAsync(c.dscal + c.save0 + q, c.dscal + slots[q], sizeof(double),
                                              hipMemcpyDeviceToDevice, c.stream));
                    saved = true;
                    later = false;
                } else {
                    bool due = false;
                    TRY(mass_correction_due(c, l, istep, due));
                    later = later || due;
                }
            }
        }
        if (saved && later)
            for (int q = 0; q < 2; q++)
                HIP_OK(hipMemcpyAsync(c.dscal + slots[q], c.dscal + c.save0 + q, sizeof(double),
                                      hipMemcpyDeviceToDevice, c.stream));
        if (reinit)
            for (const auto& ls : by_colour)
                if (!ls.empty()) TRY(launch_reinit_batch(c, ls));
    }
    // update_materials (misc.f90:79-290)
    TRY(launch_materials(c));
    return publish_volumes(c);
}

int32_t icls_ls_main(int32_t istep) {
    CTX_OR_FAIL();
    return ls_main_impl(c, istep);
}

// ---------------------------------------------------------------- Navier-Stokes
int32_t icls_average_stream_vel(double* v_avr) {
    CTX_OR_FAIL();
    TRY(launch_sum_interior(c, c.vnew, DS_V_BULK));
    // v_avr = v_avr/(itot*jtot*kmax) (misc.f90:652) on the device scalar
    if (v_avr) {
        TRY(pull_scalars(c));
        *v_avr = c.hscal[DS_V_BULK] / (c.g.nx * c.g.ny * c.g.nz);
    }
    return ICLS_OK;
}

int32_t icls_ab2(void) {
    CTX_OR_FAIL();
    TRY(launch_momentum(c));
    return launch_bc_uvw(c, c.dudt, c.dvdt, c.dwdt);
}

int32_t icls_fillps(void) {
    CTX_OR_FAIL();
    return launch_fillps(c);
}

static void rotate_jumps(Ctx& c) {
    // p_xold = p_x (gfm.f90:65-67): the new jumps overwrite the old buffers
    std::swap(c.p_x, c.p_xold);
    std::swap(c.p_y, c.p_yold);
    std::swap(c.p_z, c.p_zold);
    std::swap(c.jmask, c.jmask_old);
}

static int gfm_impl(Ctx& c, bool fuse_fillps, bool store_pjump) {
    rotate_jumps(c);
    c.p_in_pnew = false;   // fillps writes p
    const bool learned = c.learned && c.mlp.ready;

    // get_curvature, re_curvature, and for the level sets after the first their jumps
    // added inside their windows.  In the learned mode the closure takes the under-
    // resolved cells of the smoothed field: their curvature is the cell's own least-
    // squares value plus the network's correction toward the 4x finer grid's.
    auto curvature = [&](int l) -> int {
        // the first level set's call also clears the pass's total of listed cells
        if (learned) TRY(l == 0 ? launch_mlp_reset_list_total(c) : launch_mlp_reset_list(c));
        TRY(launch_curvature(c, c.lvset[l], c.curv, nullptr, nullptr, learned));
        TRY(launch_recurvature(c, c.lvset[l], c.curv, c.curv2));
        if (learned) TRY(launch_mlp_curv(c, c.lvset[l], c.curv, c.curv2));
        return l > 0 ? launch_gfm(c, l, c.curv2, false, true) : ICLS_OK;
    };
    // the first level set's jumps are written over the whole grid (they clear the
    // previous step's)
    {
        WindowScope ws(c, 0);
        TRY(curvature(0));
    }
    TRY(launch_gfm(c, 0, c.curv2, fuse_fillps && c.lmax == 1, store_pjump || c.lmax > 1));
    if (c.win.empty()) {
        for (int l = 1; l < c.lmax; l++) TRY(curvature(l));
    } else {
        // windows of one colour do not overlap: their curvatures, closures and added
        // jumps run in one launch per kernel, colour after colour
        std::vector<std::vector<int>> by_colour(c.ncolour);
        for (int l = 1; l < c.lmax; l++) by_colour[c.colour[l]].push_back(l);
        for (const auto& ls : by_colour)
            if (!ls.empty()) TRY(launch_curvature_jumps_batch(c, ls, learned));
    }
    if (fuse_fillps && c.lmax > 1) {
        TRY(launch_fillps(c));
        TRY(launch_add_pjump(c));
    }
    return ICLS_OK;
}

int32_t icls_gfm4p(void) {
    CTX_OR_FAIL();
    return gfm_impl(c, false, true);
}

int32_t icls_add_pjump(void) {
    CTX_OR_FAIL();
    return launch_add_pjump(c);
}

// solver2d on the right-hand side frhs (p itself in FP64) into p; warm inside the time
// step (the mixed refinement may start from the previous step's solution)
static int solver_impl(Ctx& c, bool warm) {
    c.p_in_pnew = false;
    return launch_solve(c, warm);
}

int32_t icls_solver2d(void) {
    CTX_OR_FAIL();
    return solver_impl(c, false);
}

int32_t icls_solver2d_cos(int32_t bc_y) {
    CTX_OR_FAIL();
    if (bc_y < ICLS_BCY_WALLS || bc_y > ICLS_BCY_OUTOUT) {
        set_error("solver2d_cos: bc_y must be 1 (Walls), 2 (InOut) or 3 (OutOut)");
        return ICLS_ERR_PARAM;
    }
    c.p_in_pnew = false;
    return launch_solve_cos(c, bc_y);
}

int32_t icls_boundp(int32_t field) {
    CTX_OR_FAIL();
    TRY(prepare_field(c, field));
    double* d = field_ptr(c, field, 0);
    if (!d || field == ICLS_F_LVSET) { set_error("bad field for boundp"); return ICLS_ERR_PARAM; }
    if (field_is_fp32(c, field)) return launch_bc_p_t_pred(c, d, nullptr);
    return launch_bc_p(c, d);
}

int32_t icls_correc(void) {
    CTX_OR_FAIL();
    return launch_correc(c);
}

int32_t icls_bounduvw(int32_t which) {
    CTX_OR_FAIL();
    if (which == 0) return launch_bc_uvw(c, c.unew, c.vnew, c.wnew);
    return launch_bc_uvw(c, c.dudt, c.dvdt, c.dwdt);
}

int32_t icls_update_pressure(int32_t first_step) {
    CTX_OR_FAIL();
    TRY(launch_pressure_norm(c));
    return launch_pressure_rotate(c, first_step);
}

int32_t icls_average_vel_at_center(void) {
    CTX_OR_FAIL();
    return launch_avg_vel(c);
}

int32_t icls_check_nan(void) {
    CTX_OR_FAIL();
    TRY(launch_nan_check(c));
    TRY(pull_scalars(c));
    if (c.hscal[DS_NAN_FLAG] != 0.) { set_error("Got NaN... Terminate brutally. :("); return ICLS_ERR_NAN; }
    return ICLS_OK;
}

// ipot (interface.f90:225-277) on request: dmin(m,n) = (dtemp(m,n) + dtemp(n,m))/2
// for m < n, zero elsewhere; lmax x lmax, column-major as the reference's dmin
int32_t icls_ipot(double* dmin) {
    CTX_OR_FAIL();
    if (!dmin) { set_error("null dmin"); return ICLS_ERR_PARAM; }
    const int L = c.lmax;
    for (int q = 0; q < L * L; q++) dmin[q] = 0.;
    if (L < 2) return ICLS_OK;
    double* dt = nullptr;
    HIP_OK(hipMalloc((void**)&dt, sizeof(double) * L * L));
    int rc = launch_ipot(c, dt);
    std::vector<double> h((size_t)L * L);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(h.data(), dt, sizeof(double) * L * L, hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) { set_error("ipot copy: %s", hipGetErrorString(e)); rc = ICLS_ERR_HIP; }
    }
    hipFree(dt);
    if (rc) return rc;
    for (int m = 0; m < L - 1; m++)
        for (int n = m + 1; n < L; n++) dmin[m + L * n] = 0.5 * (h[m + (size_t)L * n] + h[n + (size_t)L * m]);
    return ICLS_OK;
}

int32_t icls_chkdiv(int32_t istep, double* divtot, double* divmax, int32_t* ijk) {
    CTX_OR_FAIL();
    TRY(launch_chkdiv(c));
    TRY(pull_scalars(c));
    double dm = c.hscal[DS_DIV_MAX];
    long loc = (long)c.hscal[DS_DIV_MAX + 16];
    if (divtot) *divtot = c.hscal[DS_DIV_TOT];
    if (divmax) *divmax = dm;
    if (ijk) {
        if (dm > 0.) {
            ijk[0] = (int)(loc % c.g.nx) + 1;
            ijk[1] = (int)((loc / c.g.nx) % c.g.ny) + 1;
            ijk[2] = (int)(loc / ((long)c.g.nx * c.g.ny)) + 1;
        } else { ijk[0] = ijk[1] = ijk[2] = 0; }
    }
    // the reference's bound is 1e-12 (chkdiv.f90:60-86); with float velocities the
    // divergence of the stored field is at the level of float rounding, |u| 1e-7 / dx
    const double tol = c.fp32 ? 1e-3 : 1e-12;
    if (dm > tol && istep != 0) {
        set_error("Magnitude of the maximal divergence > %.0e (%.6e)", tol, dm);
        return ICLS_ERR_DIVERGENCE;
    }
    return ICLS_OK;
}

int32_t icls_chkdiv_noalloc(const void* u, const void* v, const void* w, int32_t elem_bytes, int32_t nx, int32_t ny,
                            int32_t nz, double lx, double ly, double lz, double* out, void* stream) {
    if (!u || !v || !w || !out) { set_error("null array"); return ICLS_ERR_PARAM; }
    if (elem_bytes != 4 && elem_bytes != 8) { set_error("elem_bytes must be 4 or 8, not %d", elem_bytes); return ICLS_ERR_PARAM; }
    if (nx < 1 || ny < 1 || nz < 1 || !(lx > 0.) || !(ly > 0.) || !(lz > 0.)) {
        set_error("bad grid %d x %d x %d over %g x %g x %g", nx, ny, nz, lx, ly, lz);
        return ICLS_ERR_PARAM;
    }
    Grid g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.sy = nx; g.sz = (long)nx * ny;
    g.dxi = nx / lx; g.dyi = ny / ly; g.dzi = nz / lz;   // param.f90:37-38
    g.dx = 1. / g.dxi; g.dy = 1. / g.dyi; g.dz = 1. / g.dzi;
    return chkdiv_noalloc(u, v, w, elem_bytes, g, out, (hipStream_t)stream);
}

// chkdt's combined speed (chkdt.f90:11-104) from the device maxima; dtmax = 0.5 / speed
static int chkdt_speed(Ctx& c, double& speed_out) {
    TRY(launch_chkdt_umax(c));
    TRY(pull_scalars(c));
    const Grid& g = c.g;
    const icls_params& P = c.P;
    double u_max = c.hscal[DS_UMAX], v_max = c.hscal[DS_UMAX + 1], w_max = c.hscal[DS_UMAX + 2];
    double speed_conv = u_max / g.dx + v_max / g.dy + w_max / g.dz;
    double ra = c.hscal[DS_RHO_AVR];
    double gs[6] = {fabs(P.iFr * (1. - ra / P.rho1) * P.g_unit[0]), fabs(P.iFr * (1. - ra / P.rho2) * P.g_unit[0]),
                    fabs(P.iFr * (1. - ra / P.rho1) * P.g_unit[1]), fabs(P.iFr * (1. - ra / P.rho2) * P.g_unit[1]),
                    fabs(P.iFr * (1. - ra / P.rho1) * P.g_unit[2]), fabs(P.iFr * (1. - ra / P.rho2) * P.g_unit[2])};
    double g_max = gs[0];
    for (int q = 1; q < 6; q++) if (gs[q] > g_max) g_max = gs[q];
    double h_min = fmin(g.dx, fmin(g.dy, g.dz));
    double speed_grav = sqrt(g_max / h_min);
    double v1 = fmax(P.miu1 / P.rho1, P.miu2 / P.rho2) * c.visc;
    double v2 = 2. / (g.dx * g.dx) + 2. / (g.dy * g.dy) + 2. / (g.dz * g.dz);
    double speed_visc = v1 * v2;
    double curv_max = 2. / h_min;
    double speed_surf = sqrt(curv_max / (P.We * c.rho0 * (h_min * h_min)));
    double s1 = speed_conv + speed_visc;
    speed_out = (s1 + sqrt(s1 * s1 + 4. * (speed_grav * speed_grav) + 4. * (speed_surf * speed_surf))) / 2.;
    return ICLS_OK;
}

int32_t icls_chkdt(double* dtmax) {
    CTX_OR_FAIL();
    double speed = 0.;
    TRY(chkdt_speed(c, speed));
    if (dtmax) *dtmax = 0.5 / speed;              // chkdt.f90:104
    return ICLS_OK;
}

// dt = cfl / speed (an extension: the reference keeps dt = timestep, main.f90:216);
// after the first setting dt grows at most by DT_GROWTH per update
static constexpr double DT_GROWTH = 1.2;
static int adapt_dt(Ctx& c, bool limit_growth) {
    double speed = 0.;
    TRY(chkdt_speed(c, speed));
    double dt = c.P.cfl / speed;
    if (limit_growth && dt > DT_GROWTH * c.dt) dt = DT_GROWTH * c.dt;
    if (!(dt > 0.) || !std::isfinite(dt)) { set_error("chkdt gave dt = %g", dt); return ICLS_ERR_NAN; }
    c.dt = dt;
    return ICLS_OK;
}

int32_t icls_max_vel_at_center(double* vmax) {
    CTX_OR_FAIL();
    TRY(launch_avg_vel(c));
    TRY(launch_max_vel(c));
    TRY(pull_scalars(c));
    if (vmax) *vmax = c.hscal[DS_VMAX];
    return ICLS_OK;
}

// spurious_currents (misc.f90:802-843): Ca = mu max|u_c| / sigma with the reference's
// fixed mu = 0.1 and sigma = 1; its max|u_c| is the one of max_vel_at_center
int32_t icls_spurious_currents(double* Ca) {
    double vmax = 0.;
    TRY(icls_max_vel_at_center(&vmax));
    const double mu = 0.1, sigma = 1.;
    if (Ca) *Ca = mu * vmax / sigma;
    return ICLS_OK;
}

// ---------------------------------------------------------------- init_program
static int init_derived_impl(Ctx& c) {
    if (c.learned && !c.mlp.ready) {
        // the learned mode trains its closure once, with the default options, on every
        // way into a run: a fresh start, a restart from an fld file, or uploaded fields
        icls_mlp_train_opts o;
        icls_mlp_defaults(&o);
        TRY(train_impl(c, &o));
    }
    TRY(launch_pressure_materialize(c));
    c.dt = c.P.timestep;                                   // init.f90:68
    TRY(launch_bc_uvw(c, c.unew, c.vnew, c.wnew));         // init.f90:72-76
    TRY(launch_bc_uvw(c, c.dudt, c.dvdt, c.dwdt));
    for (int l = 0; l < c.lmax; l++) TRY(launch_boundmls(c, c.lvset[l]));
    TRY(init_windows(c));
    TRY(launch_bc_p(c, c.pnew));
    TRY(launch_avg_vel(c));                                // init.f90:79
    TRY(launch_materials(c));                              // init.f90:85
    if (c.P.cfl > 0.) TRY(adapt_dt(c, false));             // needs rho_avr
    c.dt_old = c.dt;
    // init_volume = volume (init.f90:86)
    HIP_OK(hipMemcpyAsync(c.dscal + DS_COUNT_FIXED + c.lmax, c.dscal + DS_COUNT_FIXED, sizeof(double) * c.lmax,
                          hipMemcpyDeviceToDevice, c.stream));
    TRY(publish_volumes(c));
    TRY(launch_momentum_rhs_only(c));                      // init.f90:90
    // pold = pnew, p_xold = p_x ... (init.f90:107-110)
    HIP_OK(hipMemcpyAsync(c.pold, c.pnew, sizeof(double) * c.nfield, hipMemcpyDeviceToDevice, c.stream));
    TRY(launch_copy_field(c, c.p_x, c.p_xold));
    TRY(launch_copy_field(c, c.p_y, c.p_yold));
    TRY(launch_copy_field(c, c.p_z, c.p_zold));
    HIP_OK(hipMemcpyAsync(c.jmask_old, c.jmask, c.nfield, hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    c.initialised = true;
    return ICLS_OK;
}

int32_t icls_init_derived(void) {
    CTX_OR_FAIL();
    return init_derived_impl(c);
}

int32_t icls_init_program(void) {
    CTX_OR_FAIL();
    c.time = 0.;
    // init_vel_p (init.f90:148-210): fields start at zero; 'cou' sets v
    double* zf[] = {c.unew, c.vnew, c.wnew, c.dudt, c.dvdt, c.dwdt};
    for (auto* f : zf) TRY(launch_fill(c, f, 0., c.fp32));
    c.p_lazy = c.p_in_pnew = c.vbulk_fresh = false;   // pnew and pold are stored values from here
    TRY(launch_fill(c, c.pnew, 0., 0));
    TRY(launch_init_velocity(c));
    for (int l = 0; l < c.lmax; l++) TRY(launch_init_levelset(c, l, c.lvset[l]));
    return init_derived_impl(c);
}

// ---------------------------------------------------------------- checkpoint / restart
// the fields of an fld file, in file order (src/loadd.f90:76-100)
static std::vector<std::pair<int, int>> fld_fields(const Ctx& c) {
    std::vector<std::pair<int, int>> f = {{ICLS_F_UNEW, 0}, {ICLS_F_VNEW, 0}, {ICLS_F_WNEW, 0}, {ICLS_F_PNEW, 0}};
    for (int l = 0; l < c.lmax; l++) f.push_back({ICLS_F_LVSET, l});
    f.push_back({ICLS_F_JIEMIAN, 0});
    return f;
}

// between a host array of the field's Fortran shape and its interior
// (1:imax,1:jmax,1:kmax), x fastest: full -> in when to_interior, else in -> full
static void copy_interior(const Ctx& c, int field, double* full, double* in, bool to_interior) {
    int hx, hy, hz, lb;
    host_dims(c, shape_of(field), hx, hy, hz, lb);
    const int nx = c.g.nx, ny = c.g.ny, nz = c.g.nz;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t h = (size_t)(i - lb) + (size_t)hx * ((size_t)(j - lb) + (size_t)hy * (size_t)(k - lb));
                const size_t q = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
                if (to_interior) in[q] = full[h];
                else full[h] = in[q];
            }
}

static long host_count(const Ctx& c, int field) {
    int hx, hy, hz, lb;
    host_dims(c, shape_of(field), hx, hy, hz, lb);
    return (long)hx * hy * hz;
}

int32_t icls_write_fld(const char* path, int32_t nr) {
    CTX_OR_FAIL();
    if (!path) { set_error("null path"); return ICLS_ERR_PARAM; }
    const size_t n = (size_t)c.g.nx * c.g.ny * c.g.nz;
    std::vector<double> full, in(n);
    FILE* fp = fopen(path, "wb");
    if (!fp) { set_error("cannot open %s for writing", path); return ICLS_ERR_PARAM; }
    int rc = ICLS_OK;
    for (auto& fl : fld_fields(c)) {
        full.assign((size_t)host_count(c, fl.first), 0.);
        rc = icls_download(fl.first, fl.second, full.data());
        if (rc) break;
        copy_interior(c, fl.first, full.data(), in.data(), true);
        if (fwrite(in.data(), sizeof(double), n, fp) != n) { set_error("short write to %s", path); rc = ICLS_ERR_PARAM; break; }
    }
    if (!rc) {
        const double info[3] = {c.time, 1. * nr, c.dt};
        if (fwrite(info, sizeof(double), 3, fp) != 3) { set_error("short write to %s", path); rc = ICLS_ERR_PARAM; }
    }
    if (fclose(fp) != 0 && !rc) { set_error("cannot close %s", path); rc = ICLS_ERR_PARAM; }
    return rc;
}

int32_t icls_init_program_restart(const char* path, int32_t* begin) {
    CTX_OR_FAIL();
    if (!path || !begin) { set_error("null argument"); return ICLS_ERR_PARAM; }
    FILE* fp = fopen(path, "rb");
    if (!fp) { set_error("cannot open %s", path); return ICLS_ERR_PARAM; }
    const size_t n = (size_t)c.g.nx * c.g.ny * c.g.nz;
    std::vector<double> full, in(n);
    // init_vel_p (init.f90:148-210), then loadflds(0, begin)
    double* zf[] = {c.unew, c.vnew, c.wnew, c.dudt, c.dvdt, c.dwdt};
    int rc = ICLS_OK;
    for (auto* f : zf) if (!rc) rc = launch_fill(c, f, 0., c.fp32);
    c.p_lazy = c.p_in_pnew = c.vbulk_fresh = false;
    if (!rc) rc = launch_fill(c, c.pnew, 0., 0);
    if (!rc) rc = launch_init_velocity(c);
    for (auto& fl : fld_fields(c)) {
        if (rc) break;
        if (fread(in.data(), sizeof(double), n, fp) != n) { set_error("%s: file too short", path); rc = ICLS_ERR_PARAM; break; }
        if (fl.first == ICLS_F_JIEMIAN) continue;   // update_materials recomputes it
        full.assign((size_t)host_count(c, fl.first), 0.);
        if (fl.first != ICLS_F_LVSET) rc = icls_download(fl.first, fl.second, full.data());   // keep halos
        if (rc) break;
        copy_interior(c, fl.first, full.data(), in.data(), false);
        rc = icls_upload(fl.first, fl.second, full.data());
    }
    double info[3] = {0., 0., 0.};
    if (!rc && fread(info, sizeof(double), 3, fp) != 3) { set_error("%s: file too short", path); rc = ICLS_ERR_PARAM; }
    fclose(fp);
    if (rc) return rc;
    c.time = info[0];
    *begin = (int32_t)info[1];
    c.dt = info[2];
    return init_derived_impl(c);   // sets dt = timestep (init.f90:68)
}

// ---------------------------------------------------------------- output files
// The Fortran edit descriptor Ew.dE2 of the reference's writers: 0.d1..dd E+xx,
// right-justified in w columns (w asterisks when it does not fit).
static void fort_e(char* out, double v, int w, int d) {
    char m[64], f[64];
    if (v == 0.) {
        snprintf(f, sizeof f, "0.%0*dE+00", d, 0);
    } else {
        snprintf(m, sizeof m, "%.*E", d - 1, v);             // [-]D.DDDE[+-]XX
        const char* s = m;
        const bool neg = *s == '-';
        if (neg) s++;
        const char* e = strchr(s, 'E');
        const int ex = atoi(e + 1) + 1;
        char digits[40];
        int nd = 0;
        for (const char* t = s; t < e; t++)
            if (*t != '.') digits[nd++] = *t;
        digits[nd] = 0;
        snprintf(f, sizeof f, "%s0.%sE%c%02d", neg ? "-" : "", digits, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
        if ((ex < 0 ? -ex : ex) > 99) f[0] = 0;
    }
    const int n = (int)strlen(f);
    if (n == 0 || n > w) { memset(out, '*', w); out[w] = 0; return; }
    snprintf(out, w + 1, "%*s", w, f);
}

// interior (1:itot,1:jtot,1:ktot) of a downloaded field, x fastest
static int interior_of(Ctx& c, int field, int l, std::vector<double>& in) {
    std::vector<double> full((size_t)host_count(c, field));
    TRY(icls_download(field, l, full.data()));
    in.resize((size_t)c.g.nx * c.g.ny * c.g.nz);
    copy_interior(c, field, full.data(), in.data(), true);
    return ICLS_OK;
}

static bool put_be(FILE* fp, const std::vector<double>& a) {
    std::vector<uint64_t> b(a.size());
    for (size_t q = 0; q < a.size(); q++) {
        uint64_t u;
        memcpy(&u, &a[q], sizeof u);
        b[q] = __builtin_bswap64(u);
    }
    return fwrite(b.data(), sizeof(uint64_t), b.size(), fp) == b.size();
}

// vtk (write_vtk.f90:214-432) of the fields globe4vtk gathers (write_vtk.f90:45-210):
// legacy binary VTK, big-endian FP64 STRUCTURED_POINTS at the cell centres, with
// U, V, W (uccnew, vccnew, wccnew), abs(Vel), V_r = v - 6 v_bulk_desired z (1 - z)
// (v_bulk_desired = 1, param.f90:49), Pressure (pnew) and Interface (jiemian).
int32_t icls_write_vtk(const char* path) {
    CTX_OR_FAIL();
    if (!path) { set_error("null path"); return ICLS_ERR_PARAM; }
    const Grid& g = c.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    std::vector<double> u, v, w, p, jam;
    TRY(interior_of(c, ICLS_F_UCC, 0, u));
    TRY(interior_of(c, ICLS_F_VCC, 0, v));
    TRY(interior_of(c, ICLS_F_WCC, 0, w));
    TRY(interior_of(c, ICLS_F_PNEW, 0, p));
    TRY(interior_of(c, ICLS_F_JIEMIAN, 0, jam));
    FILE* fp = fopen(path, "wb");
    if (!fp) { set_error("cannot open %s for writing", path); return ICLS_ERR_PARAM; }
    std::string hdr = "# vtk DataFile Version 3.0\ntest file\nBINARY\n\nDATASET STRUCTURED_POINTS\n";
    char line[160], a[32], b[32], d[32];
    snprintf(line, sizeof line, "DIMENSIONS%12d%12d%12d\n", g.nx, g.ny, g.nz);
    hdr += line;
    fort_e(a, g.dx / 2., 14, 6); fort_e(b, g.dy / 2., 14, 6); fort_e(d, g.dz / 2., 14, 6);
    hdr += std::string("ORIGIN") + a + b + d + "\n";
    fort_e(a, g.dx, 14, 6); fort_e(b, g.dy, 14, 6); fort_e(d, g.dz, 14, 6);
    hdr += std::string("SPACING") + a + b + d + "\n\n";
    snprintf(line, sizeof line, "POINT_DATA%12lld\n", (long long)n);
    hdr += line;
    bool ok = fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size();
    auto block = [&](const char* name, const std::vector<double>& x) {
        char h[160];
        snprintf(h, sizeof h, "SCALARS %s double%12d\nLOOKUP_TABLE default\n", name, 1);
        ok = ok && fwrite(h, 1, strlen(h), fp) == strlen(h) && put_be(fp, x);
    };
    block("U", u);
    block("V", v);
    block("W", w);
    std::vector<double> t(n);
    for (size_t q = 0; q < n; q++) t[q] = sqrt(u[q] * u[q] + v[q] * v[q] + w[q] * w[q]);
    block("abs(Vel)", t);
    const double v_bulk_desired = 1.;
    for (int k = 1; k <= g.nz; k++) {
        const double coorz = (k - 0.5) * g.dz / c.P.lz;
        const double vu = 6. * v_bulk_desired * coorz * (1. - coorz);
        for (size_t q = (size_t)(k - 1) * g.nx * g.ny; q < (size_t)k * g.nx * g.ny; q++) t[q] = v[q] - vu;
    }
    block("V_r", t);
    block("Pressure", p);
    block("Interface", jam);
    ok = ok && fputc('\n', fp) != EOF;
    if (fclose(fp) != 0) ok = false;
    if (!ok) { set_error("short write to %s", path); return ICLS_ERR_PARAM; }
    return ICLS_OK;
}

// post2d (output.f90:179-427): the Tecplot ASCII mid-yz plane i = itot/2 with the
// columns y, z, u (x-centred), v and w (centred in y and z, w = 0 below k = 1),
// p, the first level set, jiemian and fixhv.  fixhv is the Heaviside of the static
// level set (init.f90:1578), which this library does not carry: its column is 0.
// The header lines are list-directed writes, spaced as flang writes them.
int32_t icls_post2d(const char* path) {
    CTX_OR_FAIL();
    if (!path) { set_error("null path"); return ICLS_ERR_PARAM; }
    const Grid& g = c.g;
    const int fid[6] = {ICLS_F_UNEW, ICLS_F_VNEW, ICLS_F_WNEW, ICLS_F_PNEW, ICLS_F_LVSET, ICLS_F_JIEMIAN};
    std::vector<double> full[6];
    int lb[6], hx[6], hy[6];
    for (int f = 0; f < 6; f++) {
        int hz;
        host_dims(c, shape_of(fid[f]), hx[f], hy[f], hz, lb[f]);
        full[f].resize((size_t)host_count(c, fid[f]));
        TRY(icls_download(fid[f], 0, full[f].data()));
    }
    auto at = [&](int f, int i, int j, int k) {
        return full[f][(size_t)(i - lb[f]) + (size_t)hx[f] * ((size_t)(j - lb[f]) + (size_t)hy[f] * (size_t)(k - lb[f]))];
    };
    FILE* fp = fopen(path, "w");
    if (!fp) { set_error("cannot open %s for writing", path); return ICLS_ERR_PARAM; }
    const int il = g.nx / 2, npoints = 0;
    bool ok = fprintf(fp, " VARIABLES = \"y\",\"z\",\"u\",\"v\",\"w\",\"p\",\"s\",\"indie\",\"fixhv\"\n") > 0;
    ok = ok && fprintf(fp, " ZONE T=\"Zone1\" I= %d  J= %d , F=POINT\n", g.ny + 2 * npoints, g.nz) > 0;
    char e[9][32];
    for (int k = 1; k <= g.nz && ok; k++) {
        for (int j = 1; j <= g.ny && ok; j++) {
            const int jm = j - 1 < 1 ? j - 1 + g.ny : j - 1;
            const double wkm = k - 1 >= 1 ? at(2, il, j, k - 1) : 0.;
            const double col[9] = {(j - 0.5) / g.dyi, (k - 0.5) / g.dzi, 0.5 * (at(0, il, j, k) + at(0, il - 1, j, k)),
                                   0.5 * (at(1, il, j, k) + at(1, il, jm, k)), 0.5 * (at(2, il, j, k) + wkm),
                                   at(3, il, j, k), at(4, il, j, k), at(5, il, j, k), 0.};
            for (int q = 0; q < 9; q++) fort_e(e[q], col[q], 15, 7);
            ok = fprintf(fp, "%s%s%s%s%s%s%s%s%s\n", e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8]) > 0;
        }
    }
    if (fclose(fp) != 0) ok = false;
    if (!ok) { set_error("short write to %s", path); return ICLS_ERR_PARAM; }
    return ICLS_OK;
}

// ---------------------------------------------------------------- one time step
int32_t icls_step(int32_t istep, int32_t begin) {
    CTX_OR_FAIL();
    if (!c.initialised) { set_error("icls_init_program has not run"); return ICLS_ERR_STATE; }
    c.time = c.time + c.dt;                                // main.f90:107
    TRY(ls_main_impl(c, istep));                           // main.f90:113
    // main.f90:124; correc already summed v unless v changed since
    if (!c.vbulk_fresh) TRY(launch_sum_interior(c, c.vnew, DS_V_BULK));
    TRY(launch_momentum(c));                               // main.f90:130
    TRY(launch_bc_uvw(c, c.dudt, c.dvdt, c.dwdt));
    TRY(gfm_impl(c, true, false));                         // main.f90:139-143
    TRY(solver_impl(c, true));                             // main.f90:149
    TRY(launch_bc_p(c, c.p));                              // main.f90:155
    TRY(launch_correc(c));                                 // main.f90:159-160
    TRY(launch_bc_uvw(c, c.unew, c.vnew, c.wnew));
    // main.f90:164-185 (norm and rotation) and main.f90:207 (NaN, a sticky flag);
    // main.f90:190: the cell-centre velocities are averaged from unew/vnew/wnew where
    // they are consumed (the next step's advection), and on download
    TRY(launch_step_end(c, istep == begin + 1));
    if (c.P.cfl > 0. && istep % c.P.ioutchk == 0) TRY(adapt_dt(c, true));   // main.f90:214
    return ICLS_OK;
}

// ---------------------------------------------------------------- profiling API
int32_t icls_profile_enable(int32_t on) {
    CTX_OR_FAIL();
    if (on && !c.prof) c.prof = new Profiler();
    if (!on && c.prof) { c.prof->flush(c.stream); delete c.prof; c.prof = nullptr; }
    return ICLS_OK;
}

int32_t icls_profile_only(const char* name) {
    CTX_OR_FAIL();
    if (!c.prof) { set_error("profiling is not enabled"); return ICLS_ERR_STATE; }
    c.prof->flush(c.stream);
    c.prof->only = name ? name : "";
    return ICLS_OK;
}

int32_t icls_profile_reset(void) {
    CTX_OR_FAIL();
    if (c.prof) { c.prof->flush(c.stream); c.prof->acc.clear(); }
    return ICLS_OK;
}

// ---------------------------------------------------------------- learned curvature closure
int32_t icls_mlp_defaults(icls_mlp_train_opts* o) {
    if (!o) return ICLS_ERR_PARAM;
    o->n_samples = 131072;
    o->epochs = 60;
    o->batch = 1024;
    o->lr = 3e-3f;
    o->r_min = 1.5f;
    o->r_max = 12.f;
    o->seed = 20261202ull;
    o->loss_first = o->loss_last = 0.f;
    return ICLS_OK;
}

// device copies of host rows, freed on scope exit
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};

static int train_impl(Ctx& c, icls_mlp_train_opts* o) {
    const long n = o->n_samples - o->n_samples % o->batch;
    if (o->batch <= 0 || o->batch % 16 || n <= 0 || o->epochs <= 0 || !(o->r_max > o->r_min) || !(o->r_min > 0.f)) {
        set_error("bad training options");
        return ICLS_ERR_PARAM;
    }
    const int nb = (int)(n / o->batch);
    DevBuf x, t, lb;
    HIP_OK(hipMalloc(&x.p, sizeof(uint16_t) * MLP_NPAD * n));
    HIP_OK(hipMalloc(&t.p, sizeof(float) * n));
    HIP_OK(hipMalloc(&lb.p, sizeof(float) * nb));   // the loss of each batch, read once per epoch
    // the cells the closure is consulted on, with a margin below the gate
    TRY(mlp_samples(c, n, o->seed, o->r_min, o->r_max, 0.75f * c.mlp.gate, (uint16_t*)x.p, (float*)t.p, nullptr));
    std::vector<float> losses((size_t)nb);
    for (int e = 0; e < o->epochs; e++) {
        // cosine decay of the learning rate to 2% over the epochs
        const float lr = o->lr * (0.02f + 0.98f * 0.5f * (1.f + cosf(3.14159265f * (float)e / (float)o->epochs)));
        for (int b = 0; b < nb; b++) {
            const long r0 = (long)b * o->batch;
            TRY(mlp_grad(c, (const uint16_t*)x.p + r0 * MLP_NPAD, (const float*)t.p + r0, o->batch, nullptr));
            HIP_OK(hipMemcpyAsync((float*)lb.p + b, c.mlp.grad + MLP_NPARAM, sizeof(float), hipMemcpyDeviceToDevice,
                                  c.stream));
            TRY(mlp_adam(c, lr));
        }
        HIP_OK(hipMemcpyAsync(losses.data(), lb.p, sizeof(float) * nb, hipMemcpyDeviceToHost, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
        double s = 0.;
        for (float v : losses) s += v;
        if (e == 0) o->loss_first = (float)(s / nb);
        o->loss_last = (float)(s / nb);
    }
    c.mlp.ready = true;
    return ICLS_OK;
}

int32_t icls_mlp_train(icls_mlp_train_opts* o) {
    CTX_OR_FAIL();
    icls_mlp_train_opts d;
    if (!o) { icls_mlp_defaults(&d); o = &d; }
    return train_impl(c, o);
}

int32_t icls_mlp_set_weights(const float* w, int32_t n) {
    CTX_OR_FAIL();
    if (!w || n != MLP_NPARAM) { set_error("expected %d weights", MLP_NPARAM); return ICLS_ERR_PARAM; }
    TRY(mlp_set_weights(c, w));
    c.mlp.ready = true;
    return ICLS_OK;
}

int32_t icls_mlp_get_weights(float* w, int32_t n) {
    CTX_OR_FAIL();
    if (!w || n != MLP_NPARAM) { set_error("expected %d weights", MLP_NPARAM); return ICLS_ERR_PARAM; }
    HIP_OK(hipMemcpyAsync(w, c.mlp.w, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return ICLS_OK;
}

int32_t icls_mlp_set_gate(float gate, float tol) {
    CTX_OR_FAIL();
    if (!(gate > 0.f) || !(tol >= 0.f)) { set_error("gate must be > 0 and tol >= 0"); return ICLS_ERR_PARAM; }
    c.mlp.gate = gate;
    c.mlp.tol = tol;
    return ICLS_OK;
}

int32_t icls_mlp_forward(const uint16_t* x, int32_t n, float* y) {
    CTX_OR_FAIL();
    if (!x || !y || n <= 0 || n % 16) { set_error("forward needs n > 0, a multiple of 16"); return ICLS_ERR_PARAM; }
    DevBuf dx, dy;
    HIP_OK(hipMalloc(&dx.p, sizeof(uint16_t) * MLP_NPAD * n));
    HIP_OK(hipMalloc(&dy.p, sizeof(float) * n));
    HIP_OK(hipMemcpyAsync(dx.p, x, sizeof(uint16_t) * MLP_NPAD * n, hipMemcpyHostToDevice, c.stream));
    TRY(mlp_forward(c, (const uint16_t*)dx.p, n, (float*)dy.p));
    HIP_OK(hipMemcpyAsync(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return ICLS_OK;
}

int32_t icls_mlp_gradient(const uint16_t* x, const float* t, int32_t n, float* grad_loss) {
    CTX_OR_FAIL();
    if (!x || !t || !grad_loss || n <= 0 || n % 16) { set_error("gradient needs n > 0, a multiple of 16"); return ICLS_ERR_PARAM; }
    DevBuf dx, dt;
    HIP_OK(hipMalloc(&dx.p, sizeof(uint16_t) * MLP_NPAD * n));
    HIP_OK(hipMalloc(&dt.p, sizeof(float) * n));
    HIP_OK(hipMemcpyAsync(dx.p, x, sizeof(uint16_t) * MLP_NPAD * n, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice, c.stream));
    TRY(mlp_grad(c, (const uint16_t*)dx.p, (const float*)dt.p, n, nullptr));
    HIP_OK(hipMemcpyAsync(grad_loss, c.mlp.grad, sizeof(float) * (MLP_NPARAM + 1), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return ICLS_OK;
}

int32_t icls_mlp_samples(int64_t n, uint64_t seed, float r_min, float r_max, uint16_t* x, float* t, float* fd) {
    CTX_OR_FAIL();
    if (!x || !t || n <= 0 || !(r_max > r_min) || !(r_min > 0.f)) { set_error("bad sample request"); return ICLS_ERR_PARAM; }
    DevBuf dx, dt, df;
    HIP_OK(hipMalloc(&dx.p, sizeof(uint16_t) * MLP_NPAD * n));
    HIP_OK(hipMalloc(&dt.p, sizeof(float) * n));
    if (fd) HIP_OK(hipMalloc(&df.p, sizeof(float) * n));
    TRY(mlp_samples(c, n, seed, r_min, r_max, 0.f, (uint16_t*)dx.p, (float*)dt.p, (float*)df.p));
    HIP_OK(hipMemcpyAsync(x, dx.p, sizeof(uint16_t) * MLP_NPAD * n, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipMemcpyAsync(t, dt.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
    if (fd) HIP_OK(hipMemcpyAsync(fd, df.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return ICLS_OK;
}

int32_t icls_mlp_listed(int32_t* n) {
    CTX_OR_FAIL();
    if (!n) return ICLS_ERR_PARAM;
    int v = 0;
    if (!c.mlp.total) { *n = 0; return ICLS_OK; }
    HIP_OK(hipMemcpyAsync(&v, c.mlp.total, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    *n = v;
    return ICLS_OK;
}

int32_t icls_kernel_times(int32_t n, char* names, double* total_ms, int64_t* counts, int32_t* n_out) {
    CTX_OR_FAIL();
    if (!c.prof) { if (n_out) *n_out = 0; return ICLS_OK; }
    c.prof->flush(c.stream);
    int q = 0;
    for (auto& kv : c.prof->acc) {
        if (q >= n) break;
        if (names) { strncpy(names + 64 * q, kv.first.c_str(), 63); names[64 * q + 63] = 0; }
        if (total_ms) total_ms[q] = kv.second.first;
        if (counts) counts[q] = kv.second.second;
        q++;
    }
    if (n_out) *n_out = q;
    return ICLS_OK;
}

}  // extern "C"
