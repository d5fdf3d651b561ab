// betse_step.hip -- MI355X (gfx950) kernels for BETSE's electrodiffusion step
// and the C ABI declared in include/betse_amd.h.
//
// One step of Simulator._run_sim_core_loop (betse/science/sim.py:1169-1366) is
// five kernels on one stream:
//
//   k_env_diffuse   per grid square, all ions: boundary reset (Q3), gradient,
//                   Nernst-Planck flux, sign-flipped boundary divergence (Q4),
//                   explicit Euler (sim.py:2209-2254)          c_env -> c_mid
//   k_mem_cell      one block per tile of whole cells (<= 256 membranes):
//                   per membrane NaK pump, GHK per ion, GHK gap junction per
//                   ion, Ca pump; per cell the CSR segment sums from LDS, the
//                   update_Co/GJ update, no_negs clamp, rho_cells -> Vmem
//                   (sim.py:1196-1283, 2086-2111, 2007-2038)   cell A -> cell B
//   k_env_exchange  per square: grid-side CSR gather of the membrane fluxes
//                   (M_divmap_mem2ecm), env update, rho_env -> surface charge
//                   (sim_toolbox.py:1191-1229, ion_current.py:75-97) c_mid -> c_env
//   k_gauss_rows    Gaussian sigma=1 along axis 0, constant-0 edges (ion_current.py:104)
//   k_gauss_cols_grad  Gaussian along axis 1, then E = -grad(screen*v_env)
//                   (ion_current.py:104-109)
//
// Floating-point expressions keep the reference's operation order (the code
// is compiled with -ffp-contract=off) so the FP64 path tracks the NumPy
// reference to a few ulps per step.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "betse_amd.h"

namespace {

constexpr int TILE_M = 256;              // membranes per k_mem_cell block
constexpr int BLOCK = 256;               // threads per grid kernel block
constexpr double FLOAT_NONCE = 1.0e-25;  // sim_toolbox.py:50

// Constants derived on the host, each in the reference's evaluation order.
struct KParams {
    int n_cells, n_mems, ny, nx, n_sq, n_ions, i_na, i_k, i_ca, pad;
    double zplus[BT_MAX_IONS];      // z + 1e-25          (electroflux zc += nonce)
    double zF[BT_MAX_IONS];         // zs * p.F           (get_current)
    double zq[BT_MAX_IONS];         // z (for alpha = (D*z*q)/(kb*T))
    double dgj_surf[BT_MAX_IONS];   // D_gj[i] * p.gj_surface
    double c_bound[BT_MAX_IONS];
    double dt, tm, gj_len, F, RT, RT_gj, q, kbT, cm_inv;
    // NaK-ATPase (sim_toolbox.py:71-122)
    double nak_coef;                // -3*block*alpha_NaK
    double k_per_na;                // -(2/3)
    double dG_RT;                   // deltaGATP/(R*T)
    double qn_scal;                 // (cADP*1e-3)*(cPi*1e-3)
    double qd_scal;                 // cATP*1e-3
    double inv_km_na, km_na, km_k, atp_km, one_p_atp_km;
    // Ca-ATPase (sim_toolbox.py:124-182)
    double ca_coef;                 // -alpha_Ca
    double qca_num;                 // cADP*cPi
    double cATP, km_ca, atp_kmca, one_p_atp_kmca;
    // environment
    double delta, two_delta, env_vol; // env_vol = cell_height*delta**2
    double surf_scale, mean_sa, kee, screen;
    double gw[5];
};

__device__ __forceinline__ bool halted(const unsigned* flags) {
    return __atomic_load_n(flags, __ATOMIC_RELAXED) != 0u;
}

// ---------------------------------------------------------------------------
// flux laws
// ---------------------------------------------------------------------------
// electroflux (sim_toolbox.py:18-69); vBA already carries its +1e-25.
template <typename FT>
__device__ __forceinline__ FT ghk(FT cA, FT cB, FT Dc, FT d, FT zc, FT vBA, FT F, FT RT) {
    FT alpha = (zc * vBA * F) / RT;
    FT exp_alpha = exp(-alpha);
    FT deno = -expm1(-alpha);
    return -((Dc * alpha) / d) * ((cB - cA * exp_alpha) / deno);
}

__device__ __forceinline__ double cube(double x) { return pow(x, 3.0); }
__device__ __forceinline__ float cube(float x) { return powf(x, 3.0f); }

// pumpNaKATP (sim_toolbox.py:71-122): returns f_Na; f_K = -(2/3)*f_Na.
template <typename FT>
__device__ __forceinline__ FT pump_nak(FT cNai, FT cNao, FT cKi, FT cKo, FT Vm, const KParams& P) {
    FT a = cNao * (FT)1e-3;
    FT b = cKi * (FT)1e-3;
    FT Qnumo = (FT)P.qn_scal * cube(a) * (b * b);
    FT c = cNai * (FT)1e-3;
    FT d = cKo * (FT)1e-3;
    FT Qdenomo = (FT)P.qd_scal * cube(c) * (d * d);
    if (Qdenomo == (FT)0) Qdenomo = (FT)1.0e-15;
    FT Q = Qnumo / Qdenomo;
    FT Keq = exp(-((FT)P.dG_RT - (((FT)P.F * Vm) / (FT)P.RT)));
    FT na = cube(cNai / (FT)P.km_na);
    FT kk = cKo / (FT)P.km_k;
    FT k2 = kk * kk;
    FT numo_E = na * k2 * (FT)P.atp_km;
    FT denomo_E = ((FT)1 + na) * ((FT)1 + k2) * (FT)P.one_p_atp_km;
    FT fwd_co = numo_E / denomo_E;
    return (FT)P.nak_coef * fwd_co * ((FT)1 - (Q / Keq));
}

// pumpCaATP (sim_toolbox.py:124-182) after no_negs on both concentrations.
template <typename FT>
__device__ __forceinline__ FT pump_ca(FT cCai, FT cCao, FT Vm, const KParams& P) {
    FT Qnumo = (FT)P.qca_num * cCao;
    FT Qdenomo = (FT)P.cATP * cCai;
    if (Qdenomo == (FT)0) Qdenomo = (FT)1.0e-16;
    FT Q = Qnumo / Qdenomo;
    FT Keq = exp(-((FT)P.dG_RT - (FT)2 * (((FT)P.F * Vm) / (FT)P.RT)));
    FT numo_E = (cCai / (FT)P.km_ca) * (FT)P.atp_kmca;
    FT denomo_E = ((FT)1 + (cCai / (FT)P.km_ca)) * (FT)P.one_p_atp_kmca;
    FT frwd = numo_E / denomo_E;
    return (FT)P.ca_coef * frwd * ((FT)1 - (Q / Keq));
}

// ---------------------------------------------------------------------------
// k_mem_cell: membrane fluxes + per-cell update + Vmem
// ---------------------------------------------------------------------------
struct MemCellArgs {
    const int32_t* mem_ptr; const int32_t* mem_cell; const int32_t* nn_mem; const int32_t* nn_cell; const int32_t* mem_sq;
    const int32_t* tile_ptr; const double* mem_sa; const double* gjopen; const double* cell_vol;
    const double* diviterm; const double* Dm;
    const double* c_env;      // pre-diffusion env (membrane + NaK gathers, Q2)
    const double* c_mid;      // post-diffusion env (Ca pump gather)
    const double* c_cell_in; const double* c_lag_in; const double* vm_in;
    double* c_cell_out; double* c_lag_out; double* vm_out;
    double* fsa;              // [n_ions][n_mems] flux_mem * mem_sa
    unsigned* flags; unsigned long long* bad_step; const long long* step_ctr;
};

template <int NI, typename FT>
__global__ void __launch_bounds__(TILE_M)
k_mem_cell(MemCellArgs A, KParams P) {
    if (halted(A.flags)) return;
    __shared__ double s_mem[NI][TILE_M];
    __shared__ double s_gj[NI][TILE_M];
    const int tile = blockIdx.x;
    const int c0 = A.tile_ptr[tile];
    const int c1 = A.tile_ptr[tile + 1];
    const int m0 = A.mem_ptr[c0];
    const int m1 = A.mem_ptr[c1];
    const int t = threadIdx.x;
    const int m = m0 + t;
    const int M = P.n_mems, C = P.n_cells, G2 = P.n_sq;

    if (m < m1) {
        const int c = A.mem_cell[m];
        const int cn = A.nn_cell[m];
        const bool bnd = (A.nn_mem[m] == m);
        const int k = A.mem_sq[m];
        const double sa = A.mem_sa[m];
        const double v = A.vm_in[c];
        const double vn = A.vm_in[cn];
        // NaK-ATPase on the lagged membrane concentrations (sim.py:1196-1234, Q1)
        FT fNa = (FT)0, fK = (FT)0;
        const bool nak = (P.i_na >= 0);
        if (nak) {
            fNa = pump_nak<FT>((FT)A.c_lag_in[P.i_na * C + c], (FT)A.c_env[P.i_na * G2 + k],
                               (FT)A.c_lag_in[P.i_k * C + c], (FT)A.c_env[P.i_k * G2 + k], (FT)v, P);
            fK = (FT)P.k_per_na * fNa;
        }
        double fm[NI];
        double vmut = v, vnmut = vn;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            double f = 0.0;                          // fluxes_mem.fill(0)
            if (nak && i == P.i_na) f = f + (double)fNa;
            if (nak && i == P.i_k) f = f + (double)fK;
            // electroflux mutates sim.vm in place: +1e-25 per ion (Q5)
            vmut = vmut + FLOAT_NONCE;
            vnmut = vnmut + FLOAT_NONCE;
            const double clag = A.c_lag_in[i * C + c];
            const FT fED = ghk<FT>((FT)A.c_env[i * G2 + k], (FT)clag, (FT)A.Dm[i * M + m], (FT)P.tm,
                                   (FT)P.zplus[i], (FT)vmut, (FT)P.F, (FT)P.RT);
            fm[i] = f + (double)fED;
            double fgj = 0.0;                        // fgj[bflags_mems] = 0 (sim.py:2203)
            if (!bnd) {
                const double vgj = vnmut - vmut;
                const double vBA = vgj + FLOAT_NONCE;
                fgj = (double)ghk<FT>((FT)clag, (FT)A.c_lag_in[i * C + cn],
                                      (FT)(P.dgj_surf[i] * A.gjopen[m]), (FT)P.gj_len,
                                      (FT)P.zplus[i], (FT)vBA, (FT)P.F, (FT)P.RT_gj);
            }
            s_gj[i][t] = (-fgj) * sa;
        }
        // Ca-ATPase after the ion loop: refreshed cell Ca, post-diffusion env Ca,
        // Vmem carrying all n_ions nonces (sim.py:2126-2159)
        double fCa = 0.0;
        if (P.i_ca >= 0) {
            double cai = A.c_cell_in[P.i_ca * C + c];
            double cao = A.c_mid[P.i_ca * G2 + k];
            if (cai < 0.0) cai = 0.0;
            if (cao < 0.0) cao = 0.0;
            fCa = (double)pump_ca<FT>((FT)cai, (FT)cao, (FT)vmut, P);
        }
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            double f = fm[i];
            if (i == P.i_ca) f = f + fCa;
            const double x = f * sa;
            s_mem[i][t] = x;
            A.fsa[i * M + m] = x;
        }
    }
    __syncthreads();

    // per-cell segment sums in membrane order (np.dot(M_sum_mems, .) / reduceat)
    for (int c = c0 + t; c < c1; c += TILE_M) {
        const int a = A.mem_ptr[c] - m0;
        const int b = A.mem_ptr[c + 1] - m0;
        const double vol = A.cell_vol[c];
        double rho = 0.0;
        bool bad = false;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            double S = s_mem[i][a];
            double Sg = s_gj[i][a];
            for (int j = a + 1; j < b; ++j) {
                S = S + s_mem[i][j];
                Sg = Sg + s_gj[i][j];
            }
            const double delta_cells = S / vol;
            const double c1v = A.c_cell_in[i * C + c] + delta_cells * P.dt;
            A.c_lag_out[i * C + c] = c1v;
            const double delta_cgj = Sg / vol;
            double c2 = c1v + P.dt * delta_cgj;
            if (c2 < 0.0) c2 = 0.0;          // no_negs clamp (Q9)
            if (isnan(c2)) bad = true;
            A.c_cell_out[i * C + c] = c2;
            rho = (i == 0) ? P.zF[0] * c2 : rho + P.zF[i] * c2;
        }
        const double vm = P.cm_inv * (rho * A.diviterm[c]);
        if (isnan(vm)) bad = true;
        A.vm_out[c] = vm;
        if (bad) {
            atomicOr(A.flags, 1u);
            atomicMin(A.bad_step, (unsigned long long)(*A.step_ctr));
        }
    }
}

// ---------------------------------------------------------------------------
// k_env_diffuse: update_ecm for every ion (sim.py:2209-2254)
// ---------------------------------------------------------------------------
struct EnvArgs {
    const double* c_env; const double* D_env; const double* Ex; const double* Ey;
    double* c_mid; double* fex; double* fey; const unsigned* flags;
};

template <typename FT>
__global__ void __launch_bounds__(BLOCK)
k_env_diffuse(EnvArgs A, KParams P) {
    if (halted(A.flags)) return;
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= P.n_sq) return;
    const int ny = P.ny, nx = P.nx;
    const int r = idx / nx, cc = idx % nx;
    const double dl = P.delta, d2 = P.two_delta;
    for (int i = 0; i < P.n_ions; ++i) {
        const double* cenv = A.c_env + (size_t)i * P.n_sq;
        const double* D = A.D_env + (size_t)i * P.n_sq;
        const double cb = P.c_bound[i];
        const double zi = P.zq[i];
        // F(r,c): state with the world boundary reset to c_env_bound (Q3)
        auto Fv = [&](int rr, int c2) -> double {
            if (rr == 0 || rr == ny - 1 || c2 == 0 || c2 == nx - 1) return cb;
            return cenv[rr * nx + c2];
        };
        auto gcx = [&](int rr, int c2) -> double {
            if (c2 == 0) return (Fv(rr, 1) - Fv(rr, 0)) / dl;
            if (c2 == nx - 1) return (Fv(rr, nx - 1) - Fv(rr, nx - 2)) / dl;
            return -(Fv(rr, c2 - 1) - Fv(rr, c2 + 1)) / d2;
        };
        auto gcy = [&](int rr, int c2) -> double {
            if (rr == 0) return (Fv(1, c2) - Fv(0, c2)) / dl;
            if (rr == ny - 1) return (Fv(ny - 1, c2) - Fv(ny - 2, c2)) / dl;
            return -(Fv(rr - 1, c2) - Fv(rr + 1, c2)) / d2;
        };
        // nernst_planck_flux with gv = -E, u = 0, mu = 0 (sim_toolbox.py:409-411)
        auto fxat = [&](int rr, int c2) -> double {
            const int kk = rr * nx + c2;
            const double Dv = D[kk];
            const double alpha = (Dv * zi * P.q) / P.kbT;
            const double gvx = -A.Ex[kk];
            return (-Dv) * gcx(rr, c2) - (alpha * gvx) * Fv(rr, c2);
        };
        auto fyat = [&](int rr, int c2) -> double {
            const int kk = rr * nx + c2;
            const double Dv = D[kk];
            const double alpha = (Dv * zi * P.q) / P.kbT;
            const double gvy = -A.Ey[kk];
            return (-Dv) * gcy(rr, c2) - (alpha * gvy) * Fv(rr, c2);
        };
        // fd.divergence(-fx, -fy): diff axis 0 / axis 1 with flipped boundary signs (Q4)
        double gx, gy;
        if (cc == 0) gx = ((-fxat(r, 0)) - (-fxat(r, 1))) / dl;
        else if (cc == nx - 1) gx = ((-fxat(r, nx - 2)) - (-fxat(r, nx - 1))) / dl;
        else gx = -((-fxat(r, cc - 1)) - (-fxat(r, cc + 1))) / d2;
        if (r == 0) gy = -((-fyat(1, cc)) - (-fyat(0, cc))) / dl;
        else if (r == ny - 1) gy = -((-fyat(ny - 1, cc)) - (-fyat(ny - 2, cc))) / dl;
        else gy = -((-fyat(r - 1, cc)) - (-fyat(r + 1, cc))) / d2;
        const double div = gx + gy;
        A.c_mid[(size_t)i * P.n_sq + idx] = Fv(r, cc) + div * P.dt;
        if (A.fex) {
            A.fex[(size_t)i * P.n_sq + idx] = fxat(r, cc);
            A.fey[(size_t)i * P.n_sq + idx] = fyat(r, cc);
        }
    }
}

// ---------------------------------------------------------------------------
// k_env_exchange: membrane -> env exchange, rho_env -> surface charge
// ---------------------------------------------------------------------------
struct ExArgs {
    const int32_t* sq_ptr; const int32_t* sq_mems; const uint8_t* has_mem;
    const double* c_mid; const double* fsa; double* c_env; double* u; const unsigned* flags;
};

__global__ void __launch_bounds__(BLOCK)
k_env_exchange(ExArgs A, KParams P) {
    if (halted(A.flags)) return;
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n_sq) return;
    const int a = A.sq_ptr[k], b = A.sq_ptr[k + 1];
    double rho = 0.0;
    for (int i = 0; i < P.n_ions; ++i) {
        double S = 0.0;
        for (int j = a; j < b; ++j) S = S + (-A.fsa[(size_t)i * P.n_mems + A.sq_mems[j]]);
        const double delta_env = S / P.env_vol;
        const double cn = A.c_mid[(size_t)i * P.n_sq + k] + delta_env * P.dt;
        A.c_env[(size_t)i * P.n_sq + k] = cn;
        rho = (i == 0) ? P.zF[0] * cn : rho + P.zF[i] * cn;
    }
    double uu = 0.0;
    if (A.has_mem[k]) uu = ((rho * P.surf_scale) / P.mean_sa) / P.kee;
    A.u[k] = uu;
}

// rho_env -> u for the current env state (used before a phase's loop)
__global__ void __launch_bounds__(BLOCK)
k_env_charge(const double* c_env, const uint8_t* has_mem, double* u, KParams P) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n_sq) return;
    double rho = 0.0;
    for (int i = 0; i < P.n_ions; ++i) {
        const double cn = c_env[(size_t)i * P.n_sq + k];
        rho = (i == 0) ? P.zF[0] * cn : rho + P.zF[i] * cn;
    }
    u[k] = has_mem[k] ? ((rho * P.surf_scale) / P.mean_sa) / P.kee : 0.0;
}

// Vmem of every cell from its concentrations (update_V cell part)
__global__ void __launch_bounds__(BLOCK)
k_cell_vm(const double* c_cell, const double* diviterm, double* vm, KParams P) {
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= P.n_cells) return;
    double rho = 0.0;
    for (int i = 0; i < P.n_ions; ++i) {
        const double x = c_cell[(size_t)i * P.n_cells + c];
        rho = (i == 0) ? P.zF[0] * x : rho + P.zF[i] * x;
    }
    vm[c] = P.cm_inv * (rho * diviterm[c]);
}

// ---------------------------------------------------------------------------
// Gaussian (scipy correlate1d symmetric order: centre, then pairs outermost
// first) along axis 0, then axis 1 + gradient.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK)
k_gauss_rows(const double* u, double* t, const unsigned* flags, KParams P) {
    if (halted(flags)) return;
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n_sq) return;
    const int nx = P.nx, ny = P.ny;
    const int r = k / nx, c = k % nx;
    double o = u[k] * P.gw[0];
    for (int j = 4; j >= 1; --j) {
        const double lo = (r - j >= 0) ? u[(r - j) * nx + c] : 0.0;
        const double hi = (r + j < ny) ? u[(r + j) * nx + c] : 0.0;
        o = o + (lo + hi) * P.gw[j];
    }
    t[k] = o;
}

__global__ void __launch_bounds__(BLOCK)
k_gauss_cols_grad(const double* t, double* v_env, double* Ex, double* Ey,
                  unsigned* flags, long long* step_ctr, int count_step, KParams P) {
    if (halted(flags)) return;
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (count_step && k == 0) step_ctr[0] = step_ctr[0] + 1;
    if (k >= P.n_sq) return;
    const int nx = P.nx, ny = P.ny;
    const int r = k / nx, c = k % nx;
    auto vat = [&](int rr, int c2) -> double {
        const double* row = t + rr * nx;
        double o = row[c2] * P.gw[0];
        for (int j = 4; j >= 1; --j) {
            const double lo = (c2 - j >= 0) ? row[c2 - j] : 0.0;
            const double hi = (c2 + j < nx) ? row[c2 + j] : 0.0;
            o = o + (lo + hi) * P.gw[j];
        }
        return o;
    };
    auto Fs = [&](int rr, int c2) -> double { return P.screen * vat(rr, c2); };
    double gx, gy;
    if (c == 0) gx = (Fs(r, 1) - Fs(r, 0)) / P.delta;
    else if (c == nx - 1) gx = (Fs(r, nx - 1) - Fs(r, nx - 2)) / P.delta;
    else gx = -(Fs(r, c - 1) - Fs(r, c + 1)) / P.two_delta;
    if (r == 0) gy = (Fs(1, c) - Fs(0, c)) / P.delta;
    else if (r == ny - 1) gy = (Fs(ny - 1, c) - Fs(ny - 2, c)) / P.delta;
    else gy = -(Fs(r - 1, c) - Fs(r + 1, c)) / P.two_delta;
    v_env[k] = vat(r, c);
    Ex[k] = -gx;
    Ey[k] = -gy;
}

}  // namespace

// ===========================================================================
// context
// ===========================================================================
struct bt_ctx {
    bt_params prm;
    bt_buffers buf;
    KParams kp;
    hipStream_t stream = nullptr;
    int device = 0;
    int n_tiles = 0;
    // workspace
    double* c_cell_b = nullptr; double* c_lag_b = nullptr; double* vm_b = nullptr;
    double* c_mid = nullptr; double* fsa = nullptr; double* u = nullptr; double* tmp = nullptr;
    unsigned* flags = nullptr; unsigned long long* bad_step = nullptr; long long* step_ctr = nullptr;
    hipGraph_t graph = nullptr; hipGraphExec_t gexec = nullptr;
    long long steps_done = 0;
    std::string err;
};

static std::string g_create_err;

static int set_err(bt_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg; else g_create_err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                          \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return set_err(ctx, BT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void derive_params(const bt_params& p, KParams& k) {
    memset(&k, 0, sizeof(k));
    k.n_cells = p.n_cells; k.n_mems = p.n_mems; k.ny = p.grid_ny; k.nx = p.grid_nx;
    k.n_sq = p.grid_ny * p.grid_nx; k.n_ions = p.n_ions;
    k.i_na = p.i_na; k.i_k = p.i_k; k.i_ca = p.i_ca;
    for (int i = 0; i < p.n_ions; ++i) {
        k.zplus[i] = p.z[i] + FLOAT_NONCE;
        k.zF[i] = p.z[i] * p.F;
        k.zq[i] = p.z[i];
        k.dgj_surf[i] = p.d_gj[i] * p.gj_surface;
        k.c_bound[i] = p.c_bound[i];
    }
    k.dt = p.dt; k.tm = p.tm; k.gj_len = p.gj_len; k.F = p.F;
    k.RT = p.R * p.T; k.RT_gj = p.R * p.T_gj; k.q = p.q; k.kbT = p.kb * p.T;
    k.cm_inv = 1.0 / p.cm;
    k.nak_coef = -3.0 * p.NaK_block * p.alpha_NaK;
    k.k_per_na = -(2.0 / 3.0);
    k.dG_RT = p.deltaGATP / (p.R * p.T);
    k.qn_scal = (p.cADP * 1e-3) * (p.cPi * 1e-3);
    k.qd_scal = p.cATP * 1e-3;
    k.km_na = p.KmNK_Na; k.km_k = p.KmNK_K;
    k.atp_km = p.cATP / p.KmNK_ATP;
    k.one_p_atp_km = 1.0 + (p.cATP / p.KmNK_ATP);
    k.ca_coef = -p.alpha_Ca;
    k.qca_num = p.cADP * p.cPi;
    k.cATP = p.cATP; k.km_ca = p.KmCa_Ca;
    k.atp_kmca = p.cATP / p.KmCa_ATP;
    k.one_p_atp_kmca = 1.0 + (p.cATP / p.KmCa_ATP);
    k.delta = p.delta; k.two_delta = 2.0 * p.delta;
    k.env_vol = p.cell_height * (p.delta * p.delta);
    k.surf_scale = p.cell_height * (p.delta * p.delta);
    k.mean_sa = p.mean_sa;
    k.kee = (p.ko_env * p.eo) * p.er;
    k.screen = (2.0 / (p.ko_env * p.delta)) * (p.cell_radius / p.true_cell_size);
    for (int j = 0; j < 5; ++j) k.gw[j] = p.gauss_w[j];
}

template <int NI>
static void launch_mem_cell_t(bt_ctx* c, const MemCellArgs& a) {
    if (c->prm.flux_mode == BT_FLUX_FP32)
        hipLaunchKernelGGL((k_mem_cell<NI, float>), dim3(c->n_tiles), dim3(TILE_M), 0, c->stream, a, c->kp);
    else
        hipLaunchKernelGGL((k_mem_cell<NI, double>), dim3(c->n_tiles), dim3(TILE_M), 0, c->stream, a, c->kp);
}

static void launch_mem_cell(bt_ctx* c, const MemCellArgs& a) {
    switch (c->prm.n_ions) {
        case 1: launch_mem_cell_t<1>(c, a); break;
        case 2: launch_mem_cell_t<2>(c, a); break;
        case 3: launch_mem_cell_t<3>(c, a); break;
        case 4: launch_mem_cell_t<4>(c, a); break;
        case 5: launch_mem_cell_t<5>(c, a); break;
        case 6: launch_mem_cell_t<6>(c, a); break;
        case 7: launch_mem_cell_t<7>(c, a); break;
        default: launch_mem_cell_t<8>(c, a); break;
    }
}

static inline int grid_blocks(int n) { return (n + BLOCK - 1) / BLOCK; }

// One step. parity 0: cell state caller -> workspace; parity 1: workspace -> caller.
static void enqueue_step(bt_ctx* c, int parity, int which = -1) {
    const bt_buffers& B = c->buf;
    const int nsq = c->kp.n_sq;
    double* ci = parity ? c->c_cell_b : B.c_cell;
    double* li = parity ? c->c_lag_b : B.c_lag;
    double* vi = parity ? c->vm_b : B.vm_cell;
    double* co = parity ? B.c_cell : c->c_cell_b;
    double* lo = parity ? B.c_lag : c->c_lag_b;
    double* vo = parity ? B.vm_cell : c->vm_b;

    if (which < 0 || which == 1) {
        EnvArgs e{B.c_env, B.D_env, B.E_x, B.E_y, c->c_mid, nullptr, nullptr, c->flags};
        if (c->prm.flux_mode == BT_FLUX_FP32)
            hipLaunchKernelGGL((k_env_diffuse<float>), dim3(grid_blocks(nsq)), dim3(BLOCK), 0, c->stream, e, c->kp);
        else
            hipLaunchKernelGGL((k_env_diffuse<double>), dim3(grid_blocks(nsq)), dim3(BLOCK), 0, c->stream, e, c->kp);
    }
    if (which < 0 || which == 0) {
        MemCellArgs a{B.mem_ptr, B.mem_cell, B.nn_mem, B.nn_cell, B.mem_sq, B.tile_ptr, B.mem_sa, B.gjopen,
                      B.cell_vol, B.diviterm, B.Dm, B.c_env, c->c_mid, ci, li, vi, co, lo, vo,
                      c->fsa, c->flags, c->bad_step, c->step_ctr};
        launch_mem_cell(c, a);
    }
    if (which < 0 || which == 2) {
        ExArgs x{B.sq_ptr, B.sq_mems, B.sq_has_mem, c->c_mid, c->fsa, B.c_env, c->u, c->flags};
        hipLaunchKernelGGL(k_env_exchange, dim3(grid_blocks(nsq)), dim3(BLOCK), 0, c->stream, x, c->kp);
    }
    if (which < 0 || which == 3)
        hipLaunchKernelGGL(k_gauss_rows, dim3(grid_blocks(nsq)), dim3(BLOCK), 0, c->stream,
                           (const double*)c->u, c->tmp, (const unsigned*)c->flags, c->kp);
    if (which < 0 || which == 4)
        hipLaunchKernelGGL(k_gauss_cols_grad, dim3(grid_blocks(nsq)), dim3(BLOCK), 0, c->stream,
                           (const double*)c->tmp, B.v_env, B.E_x, B.E_y, c->flags, c->step_ctr,
                           which < 0 ? 1 : 0, c->kp);
}

static int copy_back(bt_ctx* c) {
    const size_t nc = (size_t)c->prm.n_cells;
    const size_t ni = (size_t)c->prm.n_ions;
    HIP_TRY(c, hipMemcpyAsync(c->buf.c_cell, c->c_cell_b, ni * nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->buf.c_lag, c->c_lag_b, ni * nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->buf.vm_cell, c->vm_b, nc * 8, hipMemcpyDeviceToDevice, c->stream));
    return BT_OK;
}

extern "C" {

int bt_version(void) { return 100; }

const char* bt_last_error(const bt_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int64_t bt_steps_done(const bt_ctx* ctx) { return ctx ? ctx->steps_done : -1; }

int bt_create(bt_ctx** out, const bt_params* p, const bt_buffers* b, void* stream) {
    if (!out || !p || !b) return set_err(nullptr, BT_EINVAL, "null argument");
    *out = nullptr;
    if (p->n_ions < 1 || p->n_ions > BT_MAX_IONS) return set_err(nullptr, BT_EINVAL, "n_ions out of range");
    if (p->n_cells < 1 || p->n_mems < 1) return set_err(nullptr, BT_EINVAL, "empty cluster");
    if (p->grid_ny < 3 || p->grid_nx < 3) return set_err(nullptr, BT_EINVAL, "grid must be at least 3x3");
    if ((long long)p->grid_ny * p->grid_nx > (1LL << 30)) return set_err(nullptr, BT_EINVAL, "grid too large");
    if (p->i_na >= p->n_ions || p->i_k >= p->n_ions || p->i_ca >= p->n_ions)
        return set_err(nullptr, BT_EINVAL, "ion slot out of range");
    if ((p->i_na < 0) != (p->i_k < 0)) return set_err(nullptr, BT_EINVAL, "NaK pump needs both Na and K");
    if (p->flux_mode != BT_FLUX_FP64 && p->flux_mode != BT_FLUX_FP32)
        return set_err(nullptr, BT_EINVAL, "unknown flux_mode");
    if (b->n_tiles < 1) return set_err(nullptr, BT_EINVAL, "n_tiles < 1");
    const void* req[] = {b->mem_ptr, b->mem_cell, b->nn_mem, b->nn_cell, b->mem_sq, b->sq_ptr, b->sq_mems,
                         b->tile_ptr, b->mem_sa, b->gjopen, b->cell_vol, b->diviterm, b->Dm, b->D_env,
                         b->sq_has_mem, b->c_cell, b->c_lag, b->vm_cell, b->c_env, b->E_x, b->E_y, b->v_env};
    for (const void* q : req)
        if (!q) return set_err(nullptr, BT_EINVAL, "null device buffer");

    bt_ctx* c = new bt_ctx();
    c->prm = *p;
    c->buf = *b;
    c->n_tiles = b->n_tiles;
    c->stream = (hipStream_t)stream;
    derive_params(*p, c->kp);
    (void)hipGetDevice(&c->device);
    const size_t nc = p->n_cells, nm = p->n_mems, ni = p->n_ions, ns = (size_t)p->grid_ny * p->grid_nx;
    hipError_t e = hipSuccess;
    auto alloc = [&](void** ptr, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(ptr, bytes < 256 ? 256 : bytes);
    };
    alloc((void**)&c->c_cell_b, ni * nc * 8);
    alloc((void**)&c->c_lag_b, ni * nc * 8);
    alloc((void**)&c->vm_b, nc * 8);
    alloc((void**)&c->c_mid, ni * ns * 8);
    alloc((void**)&c->fsa, ni * nm * 8);
    alloc((void**)&c->u, ns * 8);
    alloc((void**)&c->tmp, ns * 8);
    alloc((void**)&c->flags, 256);
    alloc((void**)&c->bad_step, 256);
    alloc((void**)&c->step_ctr, 256);
    if (e == hipSuccess) e = hipMemsetAsync(c->flags, 0, 256, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->bad_step, 0xff, 256, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->step_ctr, 0, 256, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        std::string msg = std::string("workspace allocation: ") + hipGetErrorString(e);
        bt_destroy(c);
        return set_err(nullptr, BT_EHIP, msg);
    }
    *out = c;
    return BT_OK;
}

int bt_destroy(bt_ctx* c) {
    if (!c) return BT_OK;
    (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    void* ptrs[] = {c->c_cell_b, c->c_lag_b, c->vm_b, c->c_mid, c->fsa, c->u, c->tmp,
                    c->flags, c->bad_step, c->step_ctr};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete c;
    return BT_OK;
}

int bt_update_v(bt_ctx* c) {
    if (!c) return BT_EINVAL;
    const int nsq = c->kp.n_sq;
    hipLaunchKernelGGL(k_cell_vm, dim3(grid_blocks(c->kp.n_cells)), dim3(BLOCK), 0, c->stream,
                       (const double*)c->buf.c_cell, c->buf.diviterm, c->buf.vm_cell, c->kp);
    hipLaunchKernelGGL(k_env_charge, dim3(grid_blocks(nsq)), dim3(BLOCK), 0, c->stream,
                       (const double*)c->buf.c_env, c->buf.sq_has_mem, c->u, c->kp);
    enqueue_step(c, 0, 3);
    enqueue_step(c, 0, 4);
    HIP_TRY(c, hipGetLastError());
    return BT_OK;
}

int bt_step(bt_ctx* c, int32_t n_steps, int32_t use_graph) {
    if (!c) return BT_EINVAL;
    if (n_steps < 0) return set_err(c, BT_EINVAL, "n_steps < 0");
    int pairs = n_steps / 2;
    if (use_graph && pairs > 0) {
        if (!c->gexec) {
            HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(c, 0);
            enqueue_step(c, 1);
            hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
            if (e != hipSuccess) return set_err(c, BT_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
            HIP_TRY(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        }
        for (int s = 0; s < pairs; ++s) HIP_TRY(c, hipGraphLaunch(c->gexec, c->stream));
    } else {
        for (int s = 0; s < pairs; ++s) {
            enqueue_step(c, 0);
            enqueue_step(c, 1);
        }
    }
    if (n_steps & 1) {
        enqueue_step(c, 0);
        int rc = copy_back(c);
        if (rc != BT_OK) return rc;
    }
    HIP_TRY(c, hipGetLastError());
    c->steps_done += n_steps;
    return BT_OK;
}

int bt_sync(bt_ctx* c, int64_t* bad_step) {
    if (!c) return BT_EINVAL;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    unsigned flags = 0;
    unsigned long long bs = ~0ull;
    HIP_TRY(c, hipMemcpy(&flags, c->flags, sizeof(flags), hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&bs, c->bad_step, sizeof(bs), hipMemcpyDeviceToHost));
    if (bad_step) *bad_step = (bs == ~0ull) ? -1 : (int64_t)bs;
    if (flags) return set_err(c, BT_UNSTABLE, "NaN in cell concentrations or Vmem: simulation unstable");
    return BT_OK;
}

int bt_time_kernel(bt_ctx* c, int32_t kernel, int32_t n_launches, double* ms_out) {
    if (!c || !ms_out || n_launches < 1 || kernel < 0 || kernel > 4) return BT_EINVAL;
    hipEvent_t e0, e1;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    enqueue_step(c, 0, kernel);   // warm-up launch
    HIP_TRY(c, hipEventRecord(e0, c->stream));
    for (int s = 0; s < n_launches; ++s) enqueue_step(c, 0, kernel);
    HIP_TRY(c, hipEventRecord(e1, c->stream));
    HIP_TRY(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = (double)ms / n_launches;
    return BT_OK;
}

}  // extern "C"
