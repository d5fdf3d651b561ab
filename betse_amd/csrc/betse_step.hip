// betse_step.hip -- MI355X (gfx950) kernels for BETSE's electrodiffusion step
// and the C ABI declared in include/betse_amd.h.
//
// One step of Simulator._run_sim_core_loop (betse/science/sim.py:1169-1366) is
// three kernels on one stream:
//
//   k_env_field     one block per (16x16 tile of the environment grid, pair
//                   of ions) below 1024 tiles, else per tile over all ions, staged
//                   through LDS: the surface-charge field u of the previous
//                   update_V -> SciPy Gaussian (sigma 1, constant 0) -> v_env ->
//                   E = -grad(screen*v_env) (ion_current.py:93-109), then each
//                   ion's boundary reset (Q3), gradient, Nernst-Planck flux and
//                   sign-flipped boundary divergence (Q4) of update_ecm
//                   (sim.py:2209-2254); with channels, also the gating rates
//                   of every (channel, cell) from Vmem
//                                       c_env -> c_mid (squares with membranes)
//                                       c_env -> next c_env (the other squares)
//   k_mem_cell      one or two wavefronts per tile of whole cells (<= 64
//                   membranes), a lane per membrane for every ion (NaK pump,
//                   GHK, GHK gap junction, Ca pump) and voltage-gated channel
//                   (gate update from the per-cell rates, channel GHK); per
//                   cell the CSR segment sums from LDS,
//                   update_Co + GJ update, no_negs clamp, rho_cells -> Vmem
//                   (sim.py:1196-1283, 2086-2111, 2007-2038,
//                   networks.py:3115-3215)                      cell A -> cell B
//   k_env_exchange  one wavefront per ion over 64 squares (one thread per
//                   square over all ions on very large batches): the square's membrane
//                   fluxes (written by k_mem_cell in square-sorted order, the
//                   sparse form of M_divmap_mem2ecm) -> env update; then
//                   rho_env -> u (sim_toolbox.py:1191-1229, ion_current.py:75-97)
//                   over the squares with membranes only      c_mid -> next c_env
//   k_network       (when a reaction network is attached) a lane per two cells
//                   in each of four wavefronts that split the rates:
//                   the substances' explicit Euler update (networks.py:2805-2982)
//
// The env state ping-pongs between the caller's c_env and a workspace copy, like
// the cell state, since k_env_field reads neighbours that other blocks update.
//
// Floating-point expressions keep the reference's operation order (the code is
// compiled with -ffp-contract=off) so the FP64 path tracks the NumPy reference
// to a few ulps per step. The FP32-flux path evaluates the flux laws in FP32
// and trades exact operation order for speed where an FP64 ulp or an FP32 ulp
// is all it costs (reciprocal constants, one transcendental per GHK law,
// gate-update coefficients per cell); its bar is |dVmem| < 1 uV.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <string>
#include <vector>

#include "betse_amd.h"

namespace {

constexpr int TMB = BT_TILE_MEMBRANES;   // membranes per k_mem_cell tile (one wavefront)
constexpr int BLOCK = 256;               // threads per per-square kernel block
constexpr int TY = 16, TX = 16;          // k_env_field tile
// Default launch variants by problem size (bt_tuning; bt_autotune measures
// them per problem): below ENV_SPLIT_TILES env tiles k_env_field splits its
// blocks by pairs of ions (measured: pairs of ions win by 35% at 256 tiles,
// one block per tile by 10% at 1024); below EX_SPLIT_SQUARES squares with
// membranes k_env_exchange runs one wavefront per ion
constexpr int ENV_SPLIT_TILES = 1024;
constexpr int EX_SPLIT_SQUARES = 1 << 19;
constexpr double FLOAT_NONCE = 1.0e-25;  // sim_toolbox.py:50

// Constants derived on the host, each in the reference's evaluation order.
struct KParams {
    int n_cells, n_mems, n_sq, n_ions, i_na, i_k, i_ca, n_chan;
    int chan_kind[BT_MAX_CHAN];
    int chan_ion[BT_MAX_CHAN];
    int ion_chan[BT_MAX_IONS];      // channel acting on each ion, -1 if none
    double chan_max_dm[BT_MAX_CHAN];
    double chan_dt;
    double zplus[BT_MAX_IONS];      // z + 1e-25          (electroflux zc += nonce)
    double zF[BT_MAX_IONS];         // zs * p.F           (get_current)
    double zq[BT_MAX_IONS];         // z (for alpha = (D*z*q)/(kb*T))
    double dgj_surf[BT_MAX_IONS];   // D_gj[i] * p.gj_surface
    double c_bound[BT_MAX_IONS];
    double dt, tm, gj_len, F, RT, RT_gj, q, kbT, cm_inv;
    // NaK-ATPase (sim_toolbox.py:71-122)
    double k_per_na;                // -(2/3)
    double dG_RT;                   // deltaGATP/(R*T)
    double qn_scal;                 // (cADP*1e-3)*(cPi*1e-3)
    double qd_scal;                 // cATP*1e-3
    double km_na, km_k, atp_km, one_p_atp_km;
    // Ca-ATPase (sim_toolbox.py:124-182)
    double ca_coef;                 // -alpha_Ca
    double qca_num;                 // cADP*cPi
    double cATP, km_ca, atp_kmca, one_p_atp_kmca;
    double gw[5];
    // F/RT, F/RT_gj, 1/tm, 1/gj_len, 1/RT and the pumps' 1/Km rounded once,
    // to FP64 for the FP64 path and to FP32 for the FP32-flux path
    double f_rt, f_rtgj, inv_tm, inv_gjlen, inv_rt, inv_kmna, inv_kmk, inv_kmca;
    float f_rt32, f_rtgj32, inv_tm32, inv_gjlen32, inv_rt32, inv_kmna32, inv_kmk32, inv_kmca32;
};

// Per-cluster constants of a batch, derived on the host.
struct KCluster {
    int ny, nx, sq_off, pad;
    double delta, two_delta;
    double env_vol;                 // cell_height*delta**2
    double surf_scale;              // cell_height*delta**2 (rho_env_surf factor)
    double mean_sa, kee, screen;    // kee = ko_env*eo*er
    double nak_coef;                // -3*block*alpha_NaK
    // the cluster's Dm per ion and gjopen when they are one value over all of
    // its membranes (NaN otherwise): the membrane kernel then skips the loads
    double dm_u[BT_MAX_IONS];
    double gjo_u;
    // bit i: ion i has no membrane flux anywhere in the cluster (Dm zero on
    // every membrane and no pump or channel carrying it), so its membrane
    // flux rows are exact zeros that need neither writing nor summing;
    // bit PHIB_BIT: phib below is nonzero (read with the ion bits, one load)
    unsigned still_mem;
    // nonzero where an applied boundary voltage (Phi_b) reaches the cluster's
    // squares: membranes then take their Vmem less Phi_b, v_env adds it
    int phib;
    // likewise D_env per ion over the cluster's squares: k_env_field then
    // stages the constant instead of loading the array
    double denv_u[BT_MAX_IONS];
};

// KCluster::still_mem bit of clusters with an applied boundary voltage (above every ion's bit)
constexpr unsigned PHIB_BIT = 1u << 31;

// x clamped into [0, n - 1]
__device__ __forceinline__ int clampi(int x, int n) { return min(max(x, 0), n - 1); }

// Workgroups are dealt to the 8 XCDs (each with its own L2) round-robin by
// their linear index. Remap the linear index b of n so that each XCD receives
// one contiguous chunk of the work: neighbouring tiles, which share squares,
// halos and gap-junction partners, then hit the same L2.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_chunk(int b, int n) {
    const int q = n / N_XCD, r = n % N_XCD;
    const int x = b % N_XCD, k = b / N_XCD;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

__device__ __forceinline__ bool halted(const unsigned* flags) {
    return __atomic_load_n(flags, __ATOMIC_RELAXED) != 0u;
}

// Barrier for a block whose threads exchange data through LDS only: the
// fences are restricted to LDS, so the barrier does not wait for the block's
// outstanding global stores (__syncthreads fences global memory as well, and
// so waits for every store the block has issued before it).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------
// flux laws
// ---------------------------------------------------------------------------
// Quotient of computed values. The FP64 path refines the hardware reciprocal
// y of b by two Newton steps, y + y (1 - b y), then corrects the quotient
// q = a y once, q + (a - b q) y (Markstein): the IEEE quotient in all but
// rare last-bit cases for operands whose quotient and reciprocal are normal
// numbers (the laws' are), eight instructions and fewer registers than the
// scaled IEEE sequence (the membrane kernel then holds a seventh wave per
// SIMD). The FP32-flux path takes the hardware reciprocal (1 ulp) and one
// Newton step on the quotient, q + (a - b q) / b: within an ulp of the IEEE
// quotient for operands in the normal range, four instructions instead of
// the ~10 of the scaled IEEE sequence.
__device__ __forceinline__ double dv(double a, double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
}
__device__ __forceinline__ float dv(float a, float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    return fmaf(fmaf(-b, q, a), r, q);
}

// exp(x) on the FP32-flux path, for |x| < 80 (the laws' exponents stay far
// inside): the hardware exp2 (1 ulp) of t = fl(x log2 e), times 1 + r ln 2
// with r = x log2 e - t carried in FP32 (log2 e as a sum of two floats):
// within about 2 ulps of exp(x), six instructions instead of the ~11 of the
// range-reduced library expf. The FP64 path calls exp.
__device__ __forceinline__ float exp32(float x) {
    constexpr float L_HI = 1.44269502162933349609375f;     // log2(e) rounded to FP32
    constexpr float L_LO = 1.925963033500011e-08f;          // log2(e) - L_HI
    const float t = x * L_HI;
    const float r = fmaf(x, L_LO, fmaf(x, L_HI, -t));
    const float e2 = __builtin_amdgcn_exp2f(t);
    return fmaf(e2, r * 0.693147182f, e2);
}
// expm1(x) on the FP32-flux path for |x| <= 1/2: its Taylor series to x^9
// in Horner form, x + x^2 (1/2! + x (1/3! + ... + x / 9!)) (truncation below
// 1.1e-8 relative, about 1 ulp of rounding), ten instructions.
__device__ __forceinline__ float expm1_32(float x) {
    float p = 2.75573192e-6f;                 // 1/9!
    p = fmaf(p, x, 2.48015873e-5f);           // 1/8!
    p = fmaf(p, x, 1.98412698e-4f);           // 1/7!
    p = fmaf(p, x, 1.38888889e-3f);           // 1/6!
    p = fmaf(p, x, 8.33333333e-3f);           // 1/5!
    p = fmaf(p, x, 4.16666667e-2f);           // 1/4!
    p = fmaf(p, x, 1.66666667e-1f);           // 1/3!
    p = fmaf(p, x, 0.5f);                     // 1/2!
    return fmaf(p * x, x, x);
}
// a * b + c with c in an SGPR pair. A polynomial's 64-bit coefficients are
// then built by the scalar unit; left to itself the compiler builds each in
// VGPRs with two v_mov per use (wave64: four cycles each), which doubled the
// cost of the library's exp.
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// 1/k! rounded once (k! is exact in FP64 up to 18!)
constexpr double inv_fact(int k) {
    double f = 1.0;
    for (int i = 2; i <= k; ++i) f *= (double)i;
    return 1.0 / f;
}

// exp(x) on the FP64 path: x = k ln2 + r with k = rint(x log2 e) and r by two
// FMAs against ln 2 split in a 32-bit head and its remainder (Cody and Waite;
// |r| <= ln2/2 and k r exact for the laws' |x| < 700), exp(r) by its Taylor
// series to r^13 in Horner form (truncation below 5e-18 relative), then 2^k
// by ldexp. Overflow gives +inf, underflow 0, NaN NaN.
__device__ __forceinline__ double exp64(double x) {
    constexpr double LOG2E = 1.4426950408889634;
    constexpr double LN2_HI = 6.93147180369123816490e-01;   // ln 2 to 32 bits
    constexpr double LN2_LO = 1.90821492927058770002e-10;   // ln 2 - LN2_HI
    const double k = __builtin_rint(x * LOG2E);
    const double r = fma(-k, LN2_LO, fma(-k, LN2_HI, x));
    double p = fma_sc(r, inv_fact(13), inv_fact(12));
#pragma unroll
    for (int j = 11; j >= 3; --j) p = fma_sc(p, r, inv_fact(j));
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}
// expm1(x) on the FP64 path for |x| <= 1/2: x + x^2 (1/2! + x (1/3! + ... +
// x / 17!)) (truncation below 2e-21 relative).
__device__ __forceinline__ double expm1_64(double x) {
    double q = fma_sc(x, inv_fact(17), inv_fact(16));
#pragma unroll
    for (int j = 15; j >= 3; --j) q = fma_sc(q, x, inv_fact(j));
    q = fma(q, x, 0.5);
    return fma(q * x, x, x);
}

__device__ __forceinline__ float exp_path(float x) { return exp32(x); }
__device__ __forceinline__ double exp_path(double x) { return exp64(x); }

__device__ __forceinline__ float expm1_path(float x) { return expm1_32(x); }
__device__ __forceinline__ double expm1_path(double x) { return expm1_64(x); }

// electroflux (sim_toolbox.py:18-69), flux = -(Dc alpha / d) (cB - cA e^-alpha)
// / (-expm1(-alpha)) with alpha = zc vBA F / (R T); zc and vBA already carry
// their +1e-25 (Q5). The law's constant quotients are products by reciprocals
// rounded once on the host (F/RT and 1/d: an ulp each of the path's precision)
// and it evaluates one transcendental instead of two: exp(-alpha) where
// |alpha| > 1/2 (then |1 - exp(-alpha)| >= 0.39 loses nothing to
// cancellation), else expm1(-alpha) and exp(-alpha) = 1 + expm1(-alpha) (then
// in [0.6, 1.7], so the sum is within an ulp or two). FT is double on the FP64
// path, float on the FP32-flux path.
template <typename FT>
__device__ __forceinline__ void ghk_exp(FT alpha, FT& exp_alpha, FT& deno) {
    if (fabs(alpha) > (FT)0.5) {
        exp_alpha = exp_path(-alpha);
        deno = (FT)1 - exp_alpha;
    } else {
        const FT em1 = expm1_path(-alpha);
        exp_alpha = (FT)1 + em1;
        deno = -em1;
    }
}

template <typename FT>
__device__ __forceinline__ FT ghk(FT cA, FT cB, FT Dc, FT zc, FT vBA, FT f_rt, FT inv_d) {
    const FT alpha = (zc * vBA) * f_rt;
    FT exp_alpha, deno;
    ghk_exp(alpha, exp_alpha, deno);
    return -((Dc * alpha) * inv_d) * dv(cB - cA * exp_alpha, deno);
}

// ghk for a membrane's ions in turn. The exponential of the last evaluated
// exponent is kept and reused when the next ion's exponent is bit-identical
// (ions of one valence: the +1e-25 nonces that tell their voltages apart are
// lost in the rounding of any |Vmem| above ~1e-9 V), so the values are those
// ghk computes. A zero permeability gives +0 without evaluating the law (ghk
// gives a zero of either sign; the flux sums it enters start at +0, so the
// concentrations are the same bit for bit).
template <typename FT>
struct GhkLast {
    FT alpha, exp_alpha, deno;
};

template <typename FT>
__device__ __forceinline__ FT ghk_reuse(FT cA, FT cB, FT Dc, FT zc, FT vBA, FT f_rt, FT inv_d, GhkLast<FT>& L) {
    if (Dc == (FT)0) return (FT)0;
    const FT alpha = (zc * vBA) * f_rt;
    if (!(alpha == L.alpha)) {
        ghk_exp(alpha, L.exp_alpha, L.deno);
        L.alpha = alpha;
    }
    return -((Dc * alpha) * inv_d) * dv(cB - cA * L.exp_alpha, L.deno);
}

// a law constant on the path's precision: the FP64 value, or its FP32 rounding
template <typename FT>
__device__ __forceinline__ FT on_path(double c64, float c32) {
    if constexpr (sizeof(FT) == 4) return c32;
    else return c64;
}

// membrane / gap-junction / channel GHK with the path's reciprocal constants
template <typename FT>
__device__ __forceinline__ FT ghk_path(FT cA, FT cB, FT Dc, FT zc, FT vBA, double f_rt, double inv_d, float f_rt32,
                                       float inv_d32) {
    if constexpr (sizeof(FT) == 4) return ghk<float>(cA, cB, Dc, zc, vBA, f_rt32, inv_d32);
    else return ghk<double>(cA, cB, Dc, zc, vBA, f_rt, inv_d);
}

// x**3: two multiplications (NumPy calls pow; the results differ by at most an
// ulp, far inside the parity bars, and pow costs an order of magnitude more)
__device__ __forceinline__ double cube(double x) { return (x * x) * x; }
__device__ __forceinline__ float cube(float x) { return (x * x) * x; }

// x / c with c a constant of the law: the product by the reciprocal of c
// rounded once on the host (inv64 on the FP64 path, inv32 on the FP32-flux path)
template <typename FT>
__device__ __forceinline__ FT dvc(FT x, double inv64, float inv32) {
    if constexpr (sizeof(FT) == 4) return x * inv32;
    else return x * inv64;
}

// pumpNaKATP (sim_toolbox.py:71-122): returns f_Na; f_K = -(2/3)*f_Na.
template <typename FT>
__device__ __forceinline__ FT pump_nak(FT cNai, FT cNao, FT cKi, FT cKo, FT Vm, double nak_coef,
                                      const KParams& P) {
    FT a = cNao * (FT)1e-3;
    FT b = cKi * (FT)1e-3;
    FT Qnumo = (FT)P.qn_scal * cube(a) * (b * b);
    FT c = cNai * (FT)1e-3;
    FT d = cKo * (FT)1e-3;
    FT Qdenomo = (FT)P.qd_scal * cube(c) * (d * d);
    if (Qdenomo == (FT)0) Qdenomo = (FT)1.0e-15;
    FT Q = dv(Qnumo, Qdenomo);
    FT Keq = exp_path(-((FT)P.dG_RT - dvc<FT>((FT)P.F * Vm, P.inv_rt, P.inv_rt32)));
    FT na = cube(dvc<FT>(cNai, P.inv_kmna, P.inv_kmna32));
    FT kk = dvc<FT>(cKo, P.inv_kmk, P.inv_kmk32);
    FT k2 = kk * kk;
    FT numo_E = na * k2 * (FT)P.atp_km;
    FT denomo_E = ((FT)1 + na) * ((FT)1 + k2) * (FT)P.one_p_atp_km;
    FT fwd_co = dv(numo_E, denomo_E);
    return (FT)nak_coef * fwd_co * ((FT)1 - dv(Q, Keq));
}

// pumpCaATP (sim_toolbox.py:124-182) after no_negs on both concentrations.
template <typename FT>
__device__ __forceinline__ FT pump_ca(FT cCai, FT cCao, FT Vm, const KParams& P) {
    FT Qnumo = (FT)P.qca_num * cCao;
    FT Qdenomo = (FT)P.cATP * cCai;
    if (Qdenomo == (FT)0) Qdenomo = (FT)1.0e-16;
    FT Q = dv(Qnumo, Qdenomo);
    FT Keq = exp_path(-((FT)P.dG_RT - (FT)2 * dvc<FT>((FT)P.F * Vm, P.inv_rt, P.inv_rt32)));
    FT numo_E = dvc<FT>(cCai, P.inv_kmca, P.inv_kmca32) * (FT)P.atp_kmca;
    FT denomo_E = ((FT)1 + dvc<FT>(cCai, P.inv_kmca, P.inv_kmca32)) * (FT)P.one_p_atp_kmca;
    FT frwd = dv(numo_E, denomo_E);
    return (FT)P.ca_coef * frwd * ((FT)1 - dv(Q, Keq));
}

// Nav1p2._calculate_state (channels/vg_na.py:169-186); V in mV.
template <typename FT>
__device__ __forceinline__ void nav1p2_rates(FT V, FT& mInf, FT& mTau, FT& hInf, FT& hTau) {
    const FT a = V - (FT)10.0;
    const FT mAlpha = ((FT)0.182 * (a - (FT)-35.0)) / ((FT)1 - exp(-(a - (FT)-35.0) / (FT)9));
    const FT mBeta = ((FT)0.124 * (-a - (FT)35.0)) / ((FT)1 - exp(-(-a - (FT)35.0) / (FT)9));
    mInf = mAlpha / (mAlpha + mBeta);
    mTau = (FT)1 / (mAlpha + mBeta);
    hInf = (FT)1.0 / ((FT)1 + exp((V - (FT)-65.0 - (FT)10.0) / (FT)6.2));
    const FT t1 = ((FT)0.024 * (a - (FT)-50.0)) / ((FT)1 - exp(-(a - (FT)-50.0) / (FT)5));
    const FT t2 = ((FT)0.0091 * (-a - (FT)75.000123)) / ((FT)1 - exp(-(-a - (FT)75.000123) / (FT)5));
    hTau = (FT)1 / (t1 + t2);
}

// Kv1p1._calculate_state (channels/vg_k.py:167-178); V in mV.
template <typename FT>
__device__ __forceinline__ void kv1p1_rates(FT V, FT& mInf, FT& mTau, FT& hInf, FT& hTau) {
    mInf = (FT)1.0 / ((FT)1 + exp((V - (FT)-30.5) / (FT)-11.3943));
    mTau = (FT)30.0 / ((FT)1 + exp((V - (FT)-76.56) / (FT)26.1479));
    hInf = (FT)1.0 / ((FT)1 + exp((V - (FT)-30.0) / (FT)27.3943));
    hTau = (FT)15000.0 / ((FT)1 + exp((V - (FT)-160.56) / (FT)-100.0));
}

// Cav1p2._calculate_state (channels/vg_ca.py:328-340); V in mV, shifted by -10 mV.
template <typename FT>
__device__ __forceinline__ void cav1p2_rates(FT V, FT& mInf, FT& mTau, FT& hInf, FT& hTau) {
    const FT W = V - (FT)10;
    mInf = (FT)1.0 / ((FT)1 + exp((W - (FT)-30.0) / (FT)-6));
    mTau = ((FT)5.0 + (FT)20.0 / ((FT)1 + exp((W - (FT)-25.0) / (FT)5)));
    hInf = (FT)1.0 / ((FT)1 + exp((W - (FT)-80.0) / (FT)6.4));
    hTau = ((FT)20.0 + (FT)50.0 / ((FT)1 + exp((W - (FT)-40.0) / (FT)7)));
}

// Vmem carrying n sequential +1e-25 nonces (electroflux mutates sim.vm in place, Q5)
__device__ __forceinline__ double with_nonces(double v, int n) {
    for (int q = 0; q < n; ++q) v = v + FLOAT_NONCE;
    return v;
}

// steady states and time constants of one channel kind at V (mV), FP64
__device__ __forceinline__ void gate_rates(int kind, double V, double& mInf, double& mTau, double& hInf,
                                           double& hTau) {
    if (kind == BT_CHAN_NAV1P2) nav1p2_rates<double>(V, mInf, mTau, hInf, hTau);
    else if (kind == BT_CHAN_KV1P1) kv1p1_rates<double>(V, mInf, mTau, hInf, hTau);
    else cav1p2_rates<double>(V, mInf, mTau, hInf, hTau);
}

// ---------------------------------------------------------------------------
// k_mem_cell: membrane fluxes + per-cell update + Vmem
// ---------------------------------------------------------------------------
struct MemCellArgs {
    const int32_t* mem_ptr; const int32_t* mem_cell; const int32_t* nn_mem; const int32_t* nn_cell;
    const int32_t* mem_sq; const int32_t* mem_pos; const int32_t* tile_ptr; const int32_t* tile_cluster;
    const KCluster* kcl;
    const double* mem_sa; const double* gjopen; const double* cell_vol; const double* diviterm; const double* Dm;
    const double* c_env;      // pre-diffusion env (membrane + NaK gathers, Q2)
    const double* c_mid;      // post-diffusion env (Ca pump and channel gathers)
    const double* c_cell_in; const double* c_lag_in; const double* vm_in;
    double* c_cell_out; double* c_lag_out; double* vm_out;
    void* fsa;                // [n_ions][n_mems] flux_mem * mem_sa, square-sorted: FP64, or FP32 on the
                              // FP32-flux path (summed in FP64 by the exchange)
    double* gates;            // [n_chan][2][n_mems]
    const double* chan_mod;   // [n_chan][n_cells]
    const double* rates;      // [n_chan][4][n_cells] gating rates of this step (k_env_field)
    double* fch;              // [n_chan][n_mems] channel flux * mem_sa, square-sorted
    unsigned* flags; unsigned long long* bad_step; const long long* step_ctr;
    // observed step only (NULL otherwise): raw fluxes in membrane order
    double* obs_fm; double* obs_fgj; double* obs_fch; double* obs_rate;
    const double* Phi_b = nullptr;   // applied boundary voltage per square (clusters with phib)
};

// LDS of one k_mem_cell block, shared by its wavefronts
struct MemLds {
    double (*mem)[TMB];       // [NI] membrane flux * mem_sa, membrane order
    double (*gj)[TMB];        // [NI] -GJ flux * mem_sa
    double (*ch)[TMB];        // [BT_MAX_CHAN] channel flux * mem_sa
    double (*c2)[TMB];        // [NI] updated cell concentrations (WG > 1)
    int* bad;
};

// Synchronisation of the lanes that share a tile's LDS: the block (WSYNC
// false: the tile's one or two wavefronts are the block; an LDS-only barrier,
// so the tile's flux stores need not have landed) or the wavefront alone (WSYNC true: in the persistent step kernel every wavefront of a block
// works on a tile of its own; the lanes of a wavefront run in step and its LDS
// accesses complete in order, so a compiler fence is all it needs).
template <bool WSYNC>
__device__ __forceinline__ void tile_sync() {
    if constexpr (WSYNC) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        lds_barrier();
    }
}

// The work of wavefront W of a k_mem_cell block. W is a template argument so
// that every ion index below is a compile-time constant (the per-ion kernel
// parameters then load once, up front).
template <int NI, typename FT, int WG, int W, bool OBS, bool WSYNC = false>
__device__ __forceinline__ void mem_cell_wave(const MemCellArgs& A, const KParams& P, int tile, const MemLds& L) {
    constexpr int NS = (NI + WG - 1) / WG;             // ion slots per wavefront
    constexpr int NQ = (BT_MAX_CHAN + WG - 1) / WG;    // channel slots per wavefront
    const int c0 = A.tile_ptr[tile];
    const int c1 = A.tile_ptr[tile + 1];
    const int m0 = A.mem_ptr[c0];
    const int m1 = A.mem_ptr[c1];
    const int j = threadIdx.x % TMB;
    const int m = m0 + j;
    const int M = P.n_mems, C = P.n_cells, G2 = P.n_sq;
    const KCluster* KC = A.kcl + A.tile_cluster[tile];
    const double nak_coef = KC->nak_coef;
    const double gjo_u = KC->gjo_u;
    const unsigned still_mem = KC->still_mem;
    if (W == 0 && j == 0) *L.bad = 0;

    if (m < m1) {
        const int mc = A.mem_cell[m];
        const int cn = A.nn_cell[m];
        const bool bnd = (A.nn_mem[m] == m);
        const int k = A.mem_sq[m];
        const int pos = A.mem_pos[m];
        const double sa = A.mem_sa[m];
        const double gjo = (gjo_u == gjo_u) ? gjo_u : A.gjopen[m];
        // Vmem of the membrane and of its partner: (1/cm) rho_surf of the cell
        // less Phi_b at the membrane's square (update_V, sim.py:2029)
        double v = A.vm_in[mc];
        double vn = A.vm_in[cn];
        if (still_mem & PHIB_BIT) {
            v = v - A.Phi_b[k];
            vn = vn - A.Phi_b[A.mem_sq[A.nn_mem[m]]];
        }
        // the gathers of one ion slot / channel slot, issued together
        double g_env[NS], g_lag[NS], g_lagn[NS], g_dm[NS];
        auto gather_ion = [&](int s) {
            const int i = W * NS + s;
            const double du = KC->dm_u[i];
            // the cluster's uniform Dm when it has one, else the membrane's;
            // nothing for a law with zero permeability (still ion, zero
            // gap-junction diffusivity), which ghk_reuse does not evaluate
            const bool still = (still_mem >> i) & 1u;
            const bool gstill = (P.dgj_surf[i] == 0.0);
            g_env[s] = still ? 0.0 : A.c_env[i * G2 + k];
            g_lag[s] = (still && gstill) ? 0.0 : A.c_lag_in[i * C + mc];
            g_lagn[s] = gstill ? 0.0 : A.c_lag_in[i * C + cn];
            // (a wave-uniform choice: the cluster's value is already in SGPRs,
            // so a uniform Dm costs no load; a pointer chosen between the two
            // made every lane load, and measured 10% slower on the batch)
            g_dm[s] = still ? 0.0 : ((du == du) ? du : A.Dm[i * M + m]);
        };
        double h_m[NQ], h_h[NQ], h_mod[NQ], h_out[NQ], h_in[NQ];
        // the update coefficients a_m, b_m, a_h, b_h of each gate (per cell, from k_env_field)
        double r_mi[NQ], r_mt[NQ], r_hi[NQ], r_ht[NQ];
        auto gather_chan = [&](int q) {
            const int ch = W + q * WG;
            const int i = P.chan_ion[ch];
            h_m[q] = A.gates[(size_t)(2 * ch) * M + m];
            h_h[q] = A.gates[(size_t)(2 * ch + 1) * M + m];
            h_mod[q] = A.chan_mod[(size_t)ch * C + mc];
            h_out[q] = A.c_mid[i * G2 + k];
            h_in[q] = A.c_cell_in[i * C + mc];
            const double* R = A.rates + (size_t)(4 * ch) * C + mc;
            r_mi[q] = R[0];
            r_mt[q] = R[C];
            r_hi[q] = R[2 * C];
            r_ht[q] = R[3 * C];
        };
        const bool do_nak = P.i_na >= 0 && (P.i_na / NS == W || P.i_k / NS == W);
        double nai = 0.0, nao = 0.0, ki = 0.0, ko = 0.0, cai = 0.0, cao = 0.0;
        if (do_nak) {
            nai = A.c_lag_in[P.i_na * C + mc];
            nao = A.c_env[P.i_na * G2 + k];
            ki = A.c_lag_in[P.i_k * C + mc];
            ko = A.c_env[P.i_k * G2 + k];
        }
        if (P.i_ca >= 0 && P.i_ca / NS == W) {
            cai = A.c_cell_in[P.i_ca * C + mc];
            cao = A.c_mid[P.i_ca * G2 + k];
        }

        // NaK-ATPase on the lagged membrane concentrations (sim.py:1196-1234, Q1)
        FT fNa = (FT)0;
        if (do_nak) {
            fNa = pump_nak<FT>((FT)nai, (FT)nao, (FT)ki, (FT)ko, (FT)v, nak_coef, P);
            if (OBS && P.i_na / NS == W) A.obs_rate[m] = -(double)fNa;
        }
        // Vmem and the neighbour's carrying the i + 1 nonces of ion i, built up
        // slot by slot in the same sequence of additions
        double vmut = with_nonces(v, W * NS + 1);
        double vnmut = with_nonces(vn, W * NS + 1);
        // the last membrane and gap-junction exponentials (none yet: NaN
        // compares unequal to every exponent)
        GhkLast<FT> last_m{(FT)NAN, (FT)0, (FT)0}, last_g{(FT)NAN, (FT)0, (FT)0};
        // software pipeline: the gathers of slot s + 1 are issued before slot
        // s's arithmetic, so one slot's latency hides behind the other's work
        if (W * NS < NI) gather_ion(0);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int i = W * NS + s;
            if (i >= NI) break;
            if (s + 1 < NS && i + 1 < NI) gather_ion(s + 1);
            if (s > 0) {
                vmut = with_nonces(vmut, 1);
                vnmut = with_nonces(vnmut, 1);
            }
            double f = 0.0;                                  // fluxes_mem.fill(0)
            if (i == P.i_na) f = f + (double)fNa;
            if (i == P.i_k) f = f + (double)((FT)P.k_per_na * fNa);
            // GHK across the membrane and across the gap junction (sim.py:1253-1269, 2162-2206)
            const double clag = g_lag[s];
            const FT fED = ghk_reuse<FT>((FT)g_env[s], (FT)clag, (FT)g_dm[s], (FT)P.zplus[i], (FT)vmut,
                                         on_path<FT>(P.f_rt, P.f_rt32), on_path<FT>(P.inv_tm, P.inv_tm32), last_m);
            f = f + (double)fED;
            double fgj = 0.0;                                // fgj[bflags_mems] = 0 (sim.py:2203)
            if (!bnd) {
                const double vBA = (vnmut - vmut) + FLOAT_NONCE;
                fgj = (double)ghk_reuse<FT>((FT)clag, (FT)g_lagn[s], (FT)(P.dgj_surf[i] * gjo), (FT)P.zplus[i],
                                            (FT)vBA, on_path<FT>(P.f_rtgj, P.f_rtgj32),
                                            on_path<FT>(P.inv_gjlen, P.inv_gjlen32), last_g);
            }
            // Ca-ATPase after the ion loop: refreshed cell Ca, post-diffusion env Ca,
            // Vmem carrying all n_ions nonces (sim.py:2126-2159)
            if (i == P.i_ca) {
                const double cai0 = (cai < 0.0) ? 0.0 : cai;
                const double cao0 = (cao < 0.0) ? 0.0 : cao;
                f = f + (double)pump_ca<FT>((FT)cai0, (FT)cao0, (FT)with_nonces(v, NI), P);
            }
            if (OBS) {
                A.obs_fm[(size_t)i * M + m] = f;
                A.obs_fgj[(size_t)i * M + m] = fgj;
            }
            const double x = f * sa;
            L.mem[i][j] = x;
            if (!((still_mem >> i) & 1u)) {                  // the exchange reads no still row
                if constexpr (sizeof(FT) == 4) static_cast<float*>(A.fsa)[(size_t)i * M + pos] = (float)x;
                else static_cast<double*>(A.fsa)[(size_t)i * M + pos] = x;
            }
            L.gj[i][j] = (-fgj) * sa;
        }
        // voltage-gated channels (networks.py:3115-3215): gates from Vmem carrying
        // the nonces so far (their rates, per cell, come from k_env_field),
        // update_mh, open probability, GHK with the current cell and post-diffusion
        // env concentrations. Gating stays FP64 in both flux modes: the gates carry
        // state from step to step, and FP32 rates (even in cancellation-free form)
        // bias them enough to move Vmem by ~1e-7 V within a few hundred steps
        // (measured), against the 1 uV bar.
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int ch = W + q * WG;
            if (ch >= P.n_chan) break;
            gather_chan(q);
            const int i = P.chan_ion[ch];
            const double vg = with_nonces(v, NI + ch);
            const int kind = P.chan_kind[ch];
            // update_mh as x' = a x + b with the per-cell coefficients a, b of each gate
            const double mm = r_mi[q] * h_m[q] + r_mt[q];
            const double hh = r_hi[q] * h_h[q] + r_ht[q];
            A.gates[(size_t)(2 * ch) * M + m] = mm;
            A.gates[(size_t)(2 * ch + 1) * M + m] = hh;
            // m**p * h**q: p = 2 squares as NumPy does, p = 3 is cube()
            const double Pch = (kind == BT_CHAN_NAV1P2) ? cube(mm) * hh
                             : (kind == BT_CHAN_KV1P1) ? mm * (hh * hh) : (mm * mm) * hh;
            const FT Dch = (FT)(Pch * P.chan_max_dm[ch] * h_mod[q]);
            const FT fc = ghk_path<FT>((FT)h_out[q], (FT)h_in[q], Dch, (FT)P.zplus[i], (FT)(vg + FLOAT_NONCE),
                                       P.f_rt, P.inv_tm, P.f_rt32, P.inv_tm32);
            const double x = (double)fc * sa;
            L.ch[ch][j] = x;
            A.fch[(size_t)ch * M + pos] = x;
            if (OBS) A.obs_fch[(size_t)ch * M + m] = (double)fc;
        }
    }
    // per-cell inputs of the second stage, loaded once the membrane stage is
    // done: held from the start, their registers cost a wave per SIMD of
    // occupancy (measured slower on the 500k-cell batch)
    const int c = c0 + j;
    const bool own_cell = (c < c1);
    int ca = 0, cb = 0;
    double vol = 1.0, divi = 0.0;
    double cin[NS];
    if (own_cell) {
        ca = A.mem_ptr[c] - m0;
        cb = A.mem_ptr[c + 1] - m0;
        vol = A.cell_vol[c];
        divi = A.diviterm[c];
#pragma unroll
        for (int s = 0; s < NS; ++s)
            if (W * NS + s < NI) cin[s] = A.c_cell_in[(W * NS + s) * C + c];
    }

    tile_sync<WSYNC>();

    // per-cell segment sums in membrane order (np.dot(M_sum_mems, .)), then
    // update_Co, the GJ update, no_negs, and rho_cells -> Vmem (ion order)
    if (own_cell) {
        bool bad = false;
        double rho = 0.0;
        // x / vol as the product by 1/vol (an FP64 ulp; one division per cell
        // instead of three per ion)
        const double rvol = 1.0 / vol;
        auto per_vol = [&](double x) { return x * rvol; };
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int i = W * NS + s;
            if (i >= NI) break;
            double base = cin[s];
            const int ch = P.ion_chan[i];
            if (ch >= 0) {             // immediate update_Co of the channel flux
                double Sc = L.ch[ch][ca];
                for (int q = ca + 1; q < cb; ++q) Sc = Sc + L.ch[ch][q];
                base = base + per_vol(Sc) * P.dt;
            }
            double S = L.mem[i][ca];
            double Sg = L.gj[i][ca];
            for (int q = ca + 1; q < cb; ++q) {
                S = S + L.mem[i][q];
                Sg = Sg + L.gj[i][q];
            }
            const double c1v = base + per_vol(S) * P.dt;
            A.c_lag_out[i * C + c] = c1v;
            double c2 = c1v + P.dt * per_vol(Sg);
            if (c2 < 0.0) c2 = 0.0;                    // no_negs clamp (Q9)
            if (isnan(c2)) bad = true;
            A.c_cell_out[i * C + c] = c2;
            if (WG > 1) L.c2[i][j] = c2;
            else rho = (i == 0) ? P.zF[0] * c2 : rho + P.zF[i] * c2;
        }
        if (WG == 1) {
            const double vm = P.cm_inv * (rho * divi);
            if (isnan(vm)) bad = true;
            A.vm_out[c] = vm;
        }
        if (bad) *L.bad = 1;
    }
    if (WG > 1) {
        tile_sync<WSYNC>();
        if (W == 0 && own_cell) {
            double rho = P.zF[0] * L.c2[0][j];
#pragma unroll
            for (int i = 1; i < NI; ++i) rho = rho + P.zF[i] * L.c2[i][j];
            const double vm = P.cm_inv * (rho * divi);
            if (isnan(vm)) *L.bad = 1;
            A.vm_out[c] = vm;
        }
    }
    tile_sync<WSYNC>();
    if (W == 0 && j == 0 && *L.bad) {
        atomicOr(A.flags, 1u);
        atomicMin(A.bad_step, (unsigned long long)(*A.step_ctr));
    }
}

// OBS: an observed step (the raw fluxes also go to the observation buffers)
template <int NI, typename FT, int WG, bool OBS>
__global__ void __launch_bounds__(TMB * WG)
k_mem_cell(MemCellArgs A, KParams P) {
    static_assert(WG == 1 || WG == 2, "one or two wavefronts per tile");
    if (halted(A.flags)) return;
    // WG wavefronts per tile; lane j of each owns membrane m0 + j and, in the
    // per-cell stage, cell c0 + j. Wavefront w evaluates the ions
    // [w * NS, w * NS + NS) with NS = ceil(n_ions / WG) (its ion slots; Na and
    // K, the first two, share one pump evaluation) and the channels w, w + WG,
    // ..., so small problems still fill the GPU. Both wavefronts pass the same
    // barriers.
    __shared__ double s_mem[NI][TMB];
    __shared__ double s_gj[NI][TMB];
    // channel fluxes: n_chan rows, in dynamic LDS sized at launch (none
    // without channels, which leaves room for more blocks per CU)
    extern __shared__ double s_dyn[];
    double (*s_ch)[TMB] = reinterpret_cast<double (*)[TMB]>(s_dyn);
    __shared__ double s_c2[WG > 1 ? NI : 1][TMB];
    __shared__ int s_bad;
    const MemLds L{s_mem, s_gj, s_ch, s_c2, &s_bad};
    const int tile = xcd_chunk(blockIdx.x, gridDim.x);
    const int w = (WG > 1) ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / TMB)) : 0;
    if (w == 0) mem_cell_wave<NI, FT, WG, 0, OBS>(A, P, tile, L);
    else mem_cell_wave<NI, FT, WG, WG - 1, OBS>(A, P, tile, L);
}

// ---------------------------------------------------------------------------
// k_env_field: Gaussian + E field from u, then update_ecm for every ion
// ---------------------------------------------------------------------------
struct FieldArgs {
    const int* env_tiles;     // (cluster, first row, first column) of each 16x16 tile
    const KCluster* kcl;
    const double* u;          // scaled surface charge of the last update_V
    const double* c_env; const double* D_env;
    double* c_mid;            // post-diffusion env (diffuse mode)
    double* v_env; double* Ex; double* Ey;   // written in field-only mode
    const unsigned* flags;
    double* obs_fx; double* obs_fy;          // fluxes_env_x/y of the observed step, or NULL
    const double* vm;         // Vmem the step's k_mem_cell reads (diffuse mode)
    double* rates;            // [n_chan][4][n_cells] gating rates out (diffuse mode), or NULL
    // diffuse mode: a square without membranes has no exchange, so its
    // post-diffusion value is already the step's new env state and goes to
    // c_env_out; squares with membranes go to c_mid for k_env_exchange
    const uint8_t* live;      // per square: 1 if k_env_exchange updates it
    double* c_env_out;        // the step's env state buffer (never the one read here)
    const double* Phi_b = nullptr;   // applied boundary voltage per square (clusters with phib)
    // the exchange done by this kernel (k_mem_cell ran first: no ion's
    // membrane law reads post-diffusion env): each square with membranes adds
    // its membranes' fluxes (square-sorted rows of fsa over [sq_ptr[k],
    // sq_ptr[k+1])) to its diffused value, and the surface charge of the
    // step's env state goes to u_out (not u, which neighbouring tiles still
    // read); the block of tile 0 advances the step counters of ctr_mask
    const int32_t* sq_ptr = nullptr; const uint8_t* has_mem = nullptr; const void* fsa = nullptr;
    int fsa32 = 0;
    double* u_out = nullptr;
    long long* step_ctr = nullptr; unsigned ctr_mask = 0; int count_step = 0;
};

// RATES: also the gating rates (diffuse mode with channels). Divisions by the
// grid spacing and by kT are products by their FP64 reciprocals (an FP64 ulp
// each).
// the gate-update coefficients of item t_it = ch * n_cells + cell from the
// cell's Vmem vmc: the channel's rates at Vmem carrying the n_ions + ch
// nonces of the membrane stage, and update_mh (channelsabc.py:54-58),
// x' = (tau x + dt inf)/(tau + dt), as x' = a x + b with a = tau/(tau+dt),
// b = dt*inf/(tau+dt) (FP64, an ulp from the quotient form)
__device__ __forceinline__ void gate_coef_item(const FieldArgs& A, const KParams& P, int t_it, double vmc) {
    const int nc = P.n_cells;
    const int ch = t_it / nc, cell = t_it - ch * nc;
    const double V = with_nonces(vmc, P.n_ions + ch) * 1000.0 + 0.0;
    double mInf, mTau, hInf, hTau;
    gate_rates(P.chan_kind[ch], V, mInf, mTau, hInf, hTau);
    double* R = A.rates + (size_t)(4 * ch) * nc + cell;
    const double dtm = P.chan_dt;
    R[0] = mTau / (mTau + dtm);
    R[nc] = (dtm * mInf) / (mTau + dtm);
    R[2 * nc] = hTau / (hTau + dtm);
    R[3 * nc] = (dtm * hInf) / (hTau + dtm);
}


// sum of -x[q] over q in [a, b), added in order from a; the loads are issued
// eight at a time (clamped to b - 1, past-the-end values unused) so one
// square's membranes cost one memory latency instead of one each
template <typename T, int W = 8>
__device__ __forceinline__ double neg_seg_sum(const T* x, int a, int b, double S = 0.0) {
    for (int q0 = a; q0 < b; q0 += W) {
        T t[W];
#pragma unroll
        for (int u = 0; u < W; ++u) t[u] = x[min(q0 + u, b - 1)];
#pragma unroll
        for (int u = 0; u < W; ++u)
            if (q0 + u < b) S = S + (-(double)t[u]);
    }
    return S;
}

// row i of the membrane fluxes (FP64 or FP32), summed negated over [a, b) in
// FP64, continuing the sum S
template <int W = 8>
__device__ __forceinline__ double neg_flux_sum(const void* fsa, int fsa32, size_t row, int a, int b,
                                               double S = 0.0) {
    return fsa32 ? neg_seg_sum<float, W>(static_cast<const float*>(fsa) + row, a, b, S)
                 : neg_seg_sum<double, W>(static_cast<const double*>(fsa) + row, a, b, S);
}

// FUSE: the first FUSE_PF membrane fluxes of a square's row, loaded at the
// start of the ion's iteration so that their latency passes in its flux
// stages; the rest FUSE_W at a time
constexpr int FUSE_PF = 2;
constexpr int FUSE_W = 4;

// LDS of one k_env_field tile, in doubles
constexpr int ENV_UH = 6;                                        // u on tile +- 6
constexpr int ENV_A1 = (TY + 2 * ENV_UH) * (TX + 2 * ENV_UH) + (TY + 4) * (TX + 2 * ENV_UH);
constexpr int ENV_A2 = (TY + 4) * (TX + 4) + (TY + 2) * (TX + 2) + TY * (TX + 2) + (TY + 2) * TX;
constexpr int ENV_POOL = ENV_A1 > ENV_A2 ? ENV_A1 : ENV_A2;
constexpr int ENV_LDS = ENV_POOL + (TY + 4) * (TX + 4) + 2 * (TY + 2) * (TX + 2) + BT_MAX_IONS;

// One env tile's work for the ion group ion_id of n_groups (the ions
// [ion_id * ipb, ion_id * ipb + ipb), ipb = ceil(n_ions / n_groups)), with
// the block's LDS at lds (ENV_LDS doubles). With RATES the block's threads also
// evaluate the gating-rate items rate_first, rate_first + rate_stride, ...
template <bool DIFFUSE, bool RATES, bool FUSE = false>
__device__ __forceinline__ void env_field_tile(const FieldArgs& A, const KParams& P, int tile_id, int ion_id,
                                               int n_groups, double* lds, int rate_first, int rate_stride,
                                               bool check_halt = true) {
    static_assert(!FUSE || (DIFFUSE && !RATES), "the fused exchange is a diffusion step without channels");
    // the halt flag is read first and tested before the first store, so its
    // latency overlaps the block's first loads (a caller that has already
    // decided for the whole block passes check_halt = false)
    const bool stop = check_halt && halted(A.flags);
    constexpr int UH = ENV_UH, UR = TY + 2 * UH, UC = TX + 2 * UH;   // u on tile +- 6
    constexpr int TR = TY + 4;                                  // row-blurred rows tile +- 2
    constexpr int VR = TY + 4, VC = TX + 4;                     // v on tile +- 2
    constexpr int ER = TY + 2, EC = TX + 2;                     // E on tile +- 1
    // the Gaussian staging (sU, sT) is dead once v_env is built, so the
    // diffusion staging (sC, sD, sFx, sFy) reuses its LDS
    double* s_pool = lds;
    double (&sV)[VR][VC] = *reinterpret_cast<double (*)[VR][VC]>(lds + ENV_POOL);
    double (&sEx)[ER][EC] = *reinterpret_cast<double (*)[ER][EC]>(lds + ENV_POOL + VR * VC);
    double (&sEy)[ER][EC] = *reinterpret_cast<double (*)[ER][EC]>(lds + ENV_POOL + VR * VC + ER * EC);
    double* s_du = lds + ENV_POOL + VR * VC + 2 * ER * EC;
    double (&sU)[UR][UC] = *reinterpret_cast<double (*)[UR][UC]>(s_pool);
    double (&sT)[TR][UC] = *reinterpret_cast<double (*)[TR][UC]>(s_pool + UR * UC);
    double (&sC)[VR][VC] = *reinterpret_cast<double (*)[VR][VC]>(s_pool);
    double (&sD)[ER][EC] = *reinterpret_cast<double (*)[ER][EC]>(s_pool + VR * VC);
    double (&sFx)[TY][TX + 2] = *reinterpret_cast<double (*)[TY][TX + 2]>(s_pool + VR * VC + ER * EC);
    double (&sFy)[TY + 2][TX] =
        *reinterpret_cast<double (*)[TY + 2][TX]>(s_pool + VR * VC + ER * EC + TY * (TX + 2));
    const int* et = A.env_tiles + 3 * tile_id;
    const KCluster K = A.kcl[et[0]];
    const int ny = K.ny, nx = K.nx, so = K.sq_off;
    const int r0 = et[1], cc0 = et[2];
    const int tid = threadIdx.x;
    const double dl = K.delta, d2 = K.two_delta;
    const double rdl = 1.0 / dl, rd2 = 1.0 / d2;
    auto per_dl = [&](double x) { return x * rdl; };
    auto per_d2 = [&](double x) { return x * rd2; };

    // ion group ion_id takes the ions [ion_id * ipb, ion_id * ipb + ipb), with
    // ipb = ceil(n_ions / n_groups), and rebuilds its tile's field for them
    const int ipb = (P.n_ions + n_groups - 1) / n_groups;
    const int i_lo = ion_id * ipb;
    const int i_hi = min(i_lo + ipb, P.n_ions);

    // Every load the block needs first is issued here before any of them is
    // used: the gating-rate Vmem, the cluster's uniform D_env per ion, the
    // live flag, the first ion's concentrations (and D_env), then u, so their
    // latencies overlap. Later ions are loaded one ion ahead, in flight
    // through the previous ion's flux stages (barriers wait for LDS only).
    const int nc = P.n_cells;
    const int items = RATES ? P.n_chan * nc : 0;
    const int nthr = rate_stride;
    const int t_first = rate_first;
    double vm_first = 0.0;
    if (RATES && t_first >= 0 && t_first < items) vm_first = A.vm[t_first % nc];
    double du_mine = 0.0;
    if (DIFFUSE && tid < P.n_ions) du_mine = A.kcl[et[0]].denv_u[tid];

    // whether this thread's square goes to the exchange (c_mid) or is final (c_env_out)
    // (an empty asm at the store keeps the compiler from testing it, and so
    // waiting for every load issued before it, right after the load)
    unsigned live_b = 0;
    int fa = 0, fb = 0;             // FUSE: the square's membranes in fsa's rows
    unsigned hm = 0;
    if (DIFFUSE) {
        const int r = r0 + tid / TX, cc = cc0 + tid % TX;
        if (r < ny && cc < nx) {
            live_b = A.live[so + r * nx + cc];
            if (FUSE) {
                fa = A.sq_ptr[so + r * nx + cc];
                fb = A.sq_ptr[so + r * nx + cc + 1];
                hm = A.has_mem[so + r * nx + cc];
            }
        }
    }
    // The squares of ion data a thread stages: its own square of the tile
    // (row tid / TX, column tid % TX) and one square of the ring around the
    // tile, for the concentrations on tile +- 2 (the RC squares of the
    // (TY + 4) x (TX + 4) block outside the tile: thread t < RC takes ring
    // square t) and for D_env on tile +- 1 (RD squares, t < RD). Their
    // clamped in-grid offsets, LDS slots and edge flags are set once per tile,
    // so an ion's loads and staging take no index arithmetic.
    static_assert(TY == 16 && TX == 16, "the halo flux mapping assumes 16 x 16 tiles");
    constexpr int RC = VR * VC - TY * TX;
    constexpr int RD = ER * EC - TY * TX;
    // ring square t of the block of width w around the tile: the w top rows,
    // the w bottom rows, then the w left and w right squares of each tile row
    auto ring = [](int t, int w, int& sr, int& sc) {
        const int bw = TX + 2 * w;
        if (t < w * bw) {
            sr = t / bw;
            sc = t % bw;
        } else if (t < 2 * w * bw) {
            t -= w * bw;
            sr = TY + w + t / bw;
            sc = t % bw;
        } else {
            t -= 2 * w * bw;
            sr = w + t / (2 * w);
            const int q = t % (2 * w);
            sc = (q < w) ? q : TX + q;
        }
    };
    auto on_edge = [&](int r, int cc) { return r == 0 || r == ny - 1 || cc == 0 || cc == nx - 1; };
    const int lr = tid / TX, lc = tid % TX;                // own square in the tile
    const int r_own = r0 + lr, c_own = cc0 + lc;
    const bool own_in = r_own < ny && c_own < nx;
    const bool own_edge = on_edge(r_own, c_own);
    const int off_own = clampi(r_own, ny) * nx + clampi(c_own, nx);
    int crs, ccs, drs, dcs;                                // ring squares in sC and sD
    ring(min(tid, RC - 1), 2, crs, ccs);
    ring(min(tid, RD - 1), 1, drs, dcs);
    const int rc_r = r0 - 2 + crs, rc_c = cc0 - 2 + ccs;
    const int rd_r = r0 - 1 + drs, rd_c = cc0 - 1 + dcs;
    const bool cring_in = rc_r >= 0 && rc_r < ny && rc_c >= 0 && rc_c < nx;
    const bool cring_edge = on_edge(rc_r, rc_c);
    const bool dring_in = rd_r >= 0 && rd_r < ny && rd_c >= 0 && rd_c < nx;
    const int off_cring = clampi(rc_r, ny) * nx + clampi(rc_c, nx);
    const int off_dring = clampi(rd_r, ny) * nx + clampi(rd_c, nx);
    // one ion's concentrations and D_env at those squares, from clamped
    // in-grid addresses (du: the cluster's D_env of the ion when it is one
    // value, no loads; NaN when it varies)
    double tc_own = 0.0, tc_ring = 0.0, td_own = 0.0, td_ring = 0.0;
    auto load_ion = [&](int i, double du) {
        const double* cenv = A.c_env + (size_t)i * P.n_sq + so;
        const double* D = A.D_env + (size_t)i * P.n_sq + so;
        tc_own = cenv[off_own];
        tc_ring = cenv[off_cring];
        if (!(du == du)) {
            td_own = D[off_own];
            td_ring = D[off_dring];
        }
    };
    if (DIFFUSE && i_lo < i_hi) load_ion(i_lo, A.kcl[et[0]].denv_u[i_lo]);
    // u last: its loads are the only unconditional ones, so the waits for
    // them count the same on every path
    constexpr int NIT = (UR * UC + TY * TX - 1) / (TY * TX);
    double t[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int e = min(tid + it * TY * TX, UR * UC - 1);
        const int rr = r0 - UH + e / UC, cc = cc0 - UH + e % UC;
        t[it] = A.u[so + clampi(rr, ny) * nx + clampi(cc, nx)];
    }

    // the gating rates of every (channel, cell) for this step's k_mem_cell.
    // They depend on the cell's Vmem alone, so evaluating them here, one
    // (channel, cell) per thread, replaces one FP64 evaluation per membrane.
    // Run after the ion loop (the Vmem of a thread's first item was loaded
    // with the block's first loads); run first, their registers would stay
    // live beside the prefetched loads and cost occupancy.
    auto gating_rates = [&]() {
        for (int t_it = t_first; t_it >= 0 && t_it < items; t_it += nthr)
            gate_coef_item(A, P, t_it, (t_it == t_first) ? vm_first : A.vm[t_it % nc]);
    };
    if (DIFFUSE && tid < P.n_ions) s_du[tid] = du_mine;

    // u with zero padding outside the grid (gaussian_filter mode='constant')
    {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int e = tid + it * TY * TX;
            if (e < UR * UC) {
                const int rr = r0 - UH + e / UC, cc = cc0 - UH + e % UC;
                sU[e / UC][e % UC] = (rr >= 0 && rr < ny && cc >= 0 && cc < nx) ? t[it] : 0.0;
            }
        }
    }
    lds_barrier();
    // axis-0 pass (scipy correlate1d: centre tap, then pairs outermost first)
    for (int e = tid; e < TR * UC; e += TY * TX) {
        const int lr = e / UC + 4, lc = e % UC;          // row in sU
        double o = sU[lr][lc] * P.gw[0];
#pragma unroll
        for (int q = 4; q >= 1; --q) o = o + (sU[lr - q][lc] + sU[lr + q][lc]) * P.gw[q];
        sT[e / UC][lc] = o;
    }
    lds_barrier();
    // axis-1 pass -> v_env (+ Phi_b, ion_current.py:104)
    for (int e = tid; e < VR * VC; e += TY * TX) {
        const int lr = e / VC, lc = e % VC + 4;          // column in sT
        double o = sT[lr][lc] * P.gw[0];
#pragma unroll
        for (int q = 4; q >= 1; --q) o = o + (sT[lr][lc - q] + sT[lr][lc + q]) * P.gw[q];
        if (K.phib) {
            const int r = r0 - 2 + lr, cc = cc0 - 2 + e % VC;
            if (r >= 0 && r < ny && cc >= 0 && cc < nx) o = o + A.Phi_b[so + r * nx + cc];
        }
        sV[lr][e % VC] = o;
    }
    lds_barrier();
    // E = -grad(screen * v_env), one-sided at the world edges (finitediff.py:1236-1266)
    for (int e = tid; e < ER * EC; e += TY * TX) {
        const int er = e / EC, ec = e % EC;
        const int r = r0 - 1 + er, cc = cc0 - 1 + ec;
        double ex = 0.0, ey = 0.0;
        if (r >= 0 && r < ny && cc >= 0 && cc < nx) {
            const int vr = er + 1, vc = ec + 1;         // (r, cc) in sV
            double gx, gy;
            if (cc == 0) gx = per_dl((K.screen * sV[vr][vc + 1] - K.screen * sV[vr][vc]));
            else if (cc == nx - 1) gx = per_dl((K.screen * sV[vr][vc] - K.screen * sV[vr][vc - 1]));
            else gx = per_d2(-(K.screen * sV[vr][vc - 1] - K.screen * sV[vr][vc + 1]));
            if (r == 0) gy = per_dl((K.screen * sV[vr + 1][vc] - K.screen * sV[vr][vc]));
            else if (r == ny - 1) gy = per_dl((K.screen * sV[vr][vc] - K.screen * sV[vr - 1][vc]));
            else gy = per_d2(-(K.screen * sV[vr - 1][vc] - K.screen * sV[vr + 1][vc]));
            ex = -gx;
            ey = -gy;
        }
        sEx[er][ec] = ex;
        sEy[er][ec] = ey;
    }
    lds_barrier();
    // a halted run writes nothing (every thread read the same flag)
    if (stop) return;
    if (FUSE && A.count_step && tile_id == 0 && ion_id == 0 && tid == 0)
        for (int g = 0; g < BT_MAX_STREAMS; ++g)
            if (A.ctr_mask & (1u << g)) A.step_ctr[g] = A.step_ctr[g] + 1;
    if (!DIFFUSE) {
        const int r = r0 + tid / TX, cc = cc0 + tid % TX;
        if (r < ny && cc < nx) {
            A.v_env[so + r * nx + cc] = sV[tid / TX + 2][tid % TX + 2];
            A.Ex[so + r * nx + cc] = sEx[tid / TX + 1][tid % TX + 1];
            A.Ey[so + r * nx + cc] = sEy[tid / TX + 1][tid % TX + 1];
        }
        return;
    }

    double rho = 0.0;               // FUSE: the square's charge over the ions, in ion order
    for (int i = i_lo; i < i_hi; ++i) {
        const double cb = P.c_bound[i];
        const double zi = P.zq[i];
        const double zq_kt = (zi * P.q) / P.kbT;        // alpha = D * zq_kt
        lds_barrier();
        // concentrations of ion i (loaded one ion ahead) with the world
        // boundary reset to c_env_bound (Q3), zero outside the grid
        sC[lr + 2][lc + 2] = own_in ? (own_edge ? cb : tc_own) : 0.0;
        if (tid < RC) sC[crs][ccs] = cring_in ? (cring_edge ? cb : tc_ring) : 0.0;
        // D_env zero over the cluster (block-uniform): the Nernst-Planck flux
        // is a signed zero everywhere and the update keeps the
        // (boundary-reset) concentrations, so its stages are skipped. D_env
        // one value over the cluster (block-uniform) is that value, with no
        // staging; else it is staged on tile +- 1, zero outside the grid.
        const double du_i = s_du[i];
        const bool still = (du_i == 0.0);
        const bool duni = (du_i == du_i);
        // FUSE: this square's first membrane fluxes of ion i (none for a still
        // ion, whose rows are zeros, or a square without membranes)
        double pf[(FUSE && FUSE_PF > 0) ? FUSE_PF : 1];
        const bool fsum = FUSE && FUSE_PF > 0 && live_b != 0 && fa < fb && !((K.still_mem >> i) & 1u);
        if (FUSE && fsum) {
            const size_t row = (size_t)i * P.n_mems;
#pragma unroll
            for (int u = 0; u < FUSE_PF; ++u) {
                const int q = min(fa + u, fb - 1);
                pf[u] = A.fsa32 ? (double)static_cast<const float*>(A.fsa)[row + q]
                                : static_cast<const double*>(A.fsa)[row + q];
            }
        }
        if (!still && !duni) {
            sD[lr + 1][lc + 1] = own_in ? td_own : 0.0;
            if (tid < RD) sD[drs][dcs] = dring_in ? td_ring : 0.0;
        }
        // the next ion's loads, in flight through this ion's flux stages
        if (i + 1 < i_hi) load_ion(i + 1, s_du[i + 1]);
        lds_barrier();
        if (!still) {
        // Nernst-Planck fluxes (sim_toolbox.py:409-411, gv = -E, u = 0, mu = 0)
        // at the square (r, cc), which is (sr, sc) in sC and (er, ec) in sD, sEx and sEy
        auto np_flux = [&](int r, int cc, int sr, int sc, int er, int ec, double& fx, double& fy, bool want_x,
                           bool want_y) {
            const double Dv = duni ? du_i : sD[er][ec];
            const double alpha = Dv * zq_kt;
            if (want_x) {
                double g;
                if (cc == 0) g = per_dl((sC[sr][sc + 1] - sC[sr][sc]));
                else if (cc == nx - 1) g = per_dl((sC[sr][sc] - sC[sr][sc - 1]));
                else g = per_d2(-(sC[sr][sc - 1] - sC[sr][sc + 1]));
                fx = (-Dv) * g - (alpha * (-sEx[er][ec])) * sC[sr][sc];
            }
            if (want_y) {
                double g;
                if (r == 0) g = per_dl((sC[sr + 1][sc] - sC[sr][sc]));
                else if (r == ny - 1) g = per_dl((sC[sr][sc] - sC[sr - 1][sc]));
                else g = per_d2(-(sC[sr - 1][sc] - sC[sr + 1][sc]));
                fy = (-Dv) * g - (alpha * (-sEy[er][ec])) * sC[sr][sc];
            }
        };
        {
            // both fluxes at the thread's own square
            double fx = 0.0, fy = 0.0;
            if (own_in) np_flux(r_own, c_own, lr + 2, lc + 2, lr + 1, lc + 1, fx, fy, true, true);
            sFx[lr][lc + 1] = fx;
            sFy[lr + 1][lc] = fy;
        }
        if (tid < 2 * TY + 2 * TX) {
            // the halo the divergence also reads, by the first wavefront: fx
            // in the columns left and right of the tile, fy in the rows above
            // and below it
            const bool xh = tid < 2 * TY;
            const int t = xh ? tid : tid - 2 * TY;
            const int a = t % TX, side = (t / TX) ? TX + 1 : 0;     // 16 squares per side
            const int hr = xh ? a : side, hc = xh ? side : a;       // in sFy rows / sFx columns
            const int r = xh ? r0 + a : r0 - 1 + side, cc = xh ? cc0 - 1 + side : cc0 + a;
            double f = 0.0;
            if (r >= 0 && r < ny && cc >= 0 && cc < nx) {
                double fx = 0.0, fy = 0.0;
                if (xh) np_flux(r, cc, a + 2, side + 1, a + 1, side, fx, fy, true, false);
                else np_flux(r, cc, side + 1, a + 2, side, a + 1, fx, fy, false, true);
                f = xh ? fx : fy;
            }
            if (xh) sFx[hr][hc] = f;
            else sFy[hr][hc] = f;
        }
        lds_barrier();
        }
        // fd.divergence(-fx, -fy) with the sign-flipped one-sided edges (Q4), Euler step
        {
            const int lr = tid / TX, lc = tid % TX;
            const int r = r0 + lr, cc = cc0 + lc;
            if (r < ny && cc < nx) {
                double cn = sC[lr + 2][lc + 2];
                if (!still) {
                    double gx, gy;
                    if (cc == 0) gx = per_dl(((-sFx[lr][lc + 1]) - (-sFx[lr][lc + 2])));
                    else if (cc == nx - 1) gx = per_dl(((-sFx[lr][lc]) - (-sFx[lr][lc + 1])));
                    else gx = per_d2(-((-sFx[lr][lc]) - (-sFx[lr][lc + 2])));
                    if (r == 0) gy = per_dl(-((-sFy[lr + 2][lc]) - (-sFy[lr + 1][lc])));
                    else if (r == ny - 1) gy = per_dl(-((-sFy[lr + 1][lc]) - (-sFy[lr][lc])));
                    else gy = per_d2(-((-sFy[lr][lc]) - (-sFy[lr + 2][lc])));
                    const double div = gx + gy;
                    cn = cn + div * P.dt;
                }
                const size_t e = (size_t)i * P.n_sq + so + r * nx + cc;
                unsigned lv = live_b;
                asm volatile("" : "+v"(lv));
                if (FUSE) {
                    // the exchange of k_env_exchange for this square and ion:
                    // a still ion's membrane fluxes are zeros, whose sum
                    // would add a signed zero
                    if (lv != 0 && !((K.still_mem >> i) & 1u)) {
                        double S = 0.0;
                        if (fsum) {
#pragma unroll
                            for (int u = 0; u < FUSE_PF; ++u)
                                if (fa + u < fb) S = S + (-pf[u]);
                            if (fb > fa + FUSE_PF)
                                S = neg_flux_sum<FUSE_W>(A.fsa, A.fsa32, (size_t)i * P.n_mems, fa + FUSE_PF, fb, S);
                        } else {
                            S = neg_flux_sum<FUSE_W>(A.fsa, A.fsa32, (size_t)i * P.n_mems, fa, fb);
                        }
                        cn = cn + (S / K.env_vol) * P.dt;
                    }
                    A.c_env_out[e] = cn;
                    rho = (i == 0) ? P.zF[0] * cn : rho + P.zF[i] * cn;
                } else if (lv != 0) {
                    A.c_mid[e] = cn;
                } else {
                    A.c_env_out[e] = cn;
                }
                if (A.obs_fx) {
                    A.obs_fx[(size_t)i * P.n_sq + so + r * nx + cc] = still ? 0.0 : sFx[lr][lc + 1];
                    A.obs_fy[(size_t)i * P.n_sq + so + r * nx + cc] = still ? 0.0 : sFy[lr + 1][lc];
                }
            }
        }
    }
    if (FUSE) {
        const int r = r0 + tid / TX, cc = cc0 + tid % TX;
        if (r < ny && cc < nx && live_b != 0)
            A.u_out[so + r * nx + cc] = hm ? ((rho * K.surf_scale) / K.mean_sa) / K.kee : 0.0;
    }
    if (RATES) gating_rates();
}

template <bool DIFFUSE, bool RATES, bool FUSE = false>
__global__ void __launch_bounds__(TY * TX)
k_env_field(FieldArgs A, KParams P) {
    __shared__ double lds[ENV_LDS];
    // (tile, ion group) from the XCD-chunked linear block index, tiles fastest
    const int lin = xcd_chunk(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int tile_id = lin % gridDim.x, ion_id = lin / gridDim.x;
    env_field_tile<DIFFUSE, RATES, FUSE>(A, P, tile_id, ion_id, gridDim.y, lds,
                                   (blockIdx.y * gridDim.x + blockIdx.x) * (TY * TX) + threadIdx.x,
                                   gridDim.x * gridDim.y * (TY * TX));
}

// ---------------------------------------------------------------------------
// k_env_exchange: membrane -> env exchange, rho_env -> surface charge
// ---------------------------------------------------------------------------
struct ExArgs {
    const int32_t* sq_ptr; const uint8_t* has_mem; const int32_t* sq_cluster; const KCluster* kcl;
    const double* c_mid; const void* fsa; const double* fch; double* c_env; double* u;
    int fsa32;                // fsa holds FP32 values (the FP32-flux path)
    // step counters: block 0 of a counted step advances step_ctr[g] for each
    // bit g of ctr_mask (one counter per stream group; all of them when the
    // launch covers the whole problem)
    const unsigned* flags; long long* step_ctr; int count_step; unsigned ctr_mask;
    // the squares the exchange updates (membranes or a surface charge), ascending;
    // the others keep u = 0 and take their env state from k_env_field
    const int32_t* sq_live; int n_live;
};

constexpr int EXQ = 64;

__device__ __forceinline__ void advance_step(const ExArgs& A) {
#pragma unroll
    for (int g = 0; g < BT_MAX_STREAMS; ++g)
        if (A.ctr_mask & (1u << g)) A.step_ctr[g] = A.step_ctr[g] + 1;
}

// One wavefront per ion over EXQ squares; the charge of each square is then
// summed over the ions in ion order from LDS.

__global__ void __launch_bounds__(BT_MAX_IONS * EXQ)
k_env_exchange(ExArgs A, KParams P) {
    if (halted(A.flags)) return;
    __shared__ double s_cn[BT_MAX_IONS][EXQ];
    const int i = threadIdx.x / EXQ, j = threadIdx.x % EXQ;
    const int q = blockIdx.x * EXQ + j;
    if (A.count_step && blockIdx.x == 0 && threadIdx.x == 0) advance_step(A);
    const bool live = (q < A.n_live);
    const int k = live ? A.sq_live[q] : 0;
    double surf_scale = 0.0, mean_sa = 1.0, kee = 1.0;
    bool hm = false;
    if (live) {
        const int a = A.sq_ptr[k], b = A.sq_ptr[k + 1];
        const KCluster* K = A.kcl + A.sq_cluster[k];
        const double env_vol = K->env_vol;
        if (i == 0) {          // the surface charge's inputs, fetched with the rest
            surf_scale = K->surf_scale;
            mean_sa = K->mean_sa;
            kee = K->kee;
            hm = A.has_mem[k] != 0;
        }
        double cn = A.c_mid[(size_t)i * P.n_sq + k];
        const int ch = P.ion_chan[i];
        if (ch >= 0) {
            const double Sc = neg_seg_sum(A.fch + (size_t)ch * P.n_mems, a, b);
            cn = cn + (Sc / env_vol) * P.dt;
        }
        if (!((K->still_mem >> i) & 1u)) {         // a still ion's fluxes are zeros
            const double S = neg_flux_sum(A.fsa, A.fsa32, (size_t)i * P.n_mems, a, b);
            cn = cn + (S / env_vol) * P.dt;
        }
        A.c_env[(size_t)i * P.n_sq + k] = cn;
        s_cn[i][j] = cn;
    }
    lds_barrier();
    if (i == 0 && live) {
        double rho = P.zF[0] * s_cn[0][j];
        for (int ii = 1; ii < P.n_ions; ++ii) rho = rho + P.zF[ii] * s_cn[ii][j];
        double uu = 0.0;
        if (hm) uu = ((rho * surf_scale) / mean_sa) / kee;
        A.u[k] = uu;
    }
}

// the exchange of the q-th square with membranes, over all ions
__device__ __forceinline__ void exchange_square(const ExArgs& A, const KParams& P, int q) {
    const int k = A.sq_live[q];
    const int a = A.sq_ptr[k], b = A.sq_ptr[k + 1];
    const KCluster& K = A.kcl[A.sq_cluster[k]];
    double rho = 0.0;
    for (int i = 0; i < P.n_ions; ++i) {
        double cn = A.c_mid[(size_t)i * P.n_sq + k];
        const int ch = P.ion_chan[i];
        if (ch >= 0) {
            const double Sc = neg_seg_sum(A.fch + (size_t)ch * P.n_mems, a, b);
            cn = cn + (Sc / K.env_vol) * P.dt;
        }
        // a still ion's membrane fluxes are zeros: their sum would add a
        // signed zero, which leaves cn as it is
        if (!((K.still_mem >> i) & 1u)) {
            const double S = neg_flux_sum(A.fsa, A.fsa32, (size_t)i * P.n_mems, a, b);
            cn = cn + (S / K.env_vol) * P.dt;
        }
        A.c_env[(size_t)i * P.n_sq + k] = cn;
        rho = (i == 0) ? P.zF[0] * cn : rho + P.zF[i] * cn;
    }
    double uu = 0.0;
    if (A.has_mem[k]) uu = ((rho * K.surf_scale) / K.mean_sa) / K.kee;
    A.u[k] = uu;
}

// one thread per square over all ions (large batches: enough squares to fill the GPU)
__global__ void __launch_bounds__(BLOCK)
k_env_exchange_sq(ExArgs A, KParams P) {
    if (halted(A.flags)) return;
    const int q = blockIdx.x * BLOCK + threadIdx.x;
    if (A.count_step && q == 0) advance_step(A);
    if (q >= A.n_live) return;
    exchange_square(A, P, q);
}

// rho_env -> u for the current env state (before a phase's loop)
__global__ void __launch_bounds__(BLOCK)
k_env_charge(const double* c_env, const uint8_t* has_mem, const int32_t* sq_cluster, const KCluster* kcl,
             double* u, KParams P) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n_sq) return;
    double rho = 0.0;
    for (int i = 0; i < P.n_ions; ++i) {
        const double cn = c_env[(size_t)i * P.n_sq + k];
        rho = (i == 0) ? P.zF[0] * cn : rho + P.zF[i] * cn;
    }
    const KCluster& K = kcl[sq_cluster[k]];
    u[k] = has_mem[k] ? ((rho * K.surf_scale) / K.mean_sa) / K.kee : 0.0;
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
// per-cell reaction network (MasterOfNetworks.run_loop, networks.py:2805-2982)
// ---------------------------------------------------------------------------

// The network as the kernel reads it, built on the host from bt_network:
// quotients by constants as products by reciprocals (1/Km, 1/gad_scale,
// 1/Keq, each an ulp from the quotient), and per substance the list of its
// (rate, coefficient) terms, the nonzero entries of its row of the reaction
// matrix in ascending rate order (the order of np.dot(reaction_matrix, rates)).
struct KNetRate {
    int kind, species, reversible, n_act, n_inh, n_reac, n_prod, pad;
    int act[BT_NET_MAX_HILL], inh[BT_NET_MAX_HILL], reac[BT_NET_MAX_TERMS], prod[BT_NET_MAX_TERMS];
    double act_rkm[BT_NET_MAX_HILL], act_n[BT_NET_MAX_HILL], inh_rkm[BT_NET_MAX_HILL], inh_n[BT_NET_MAX_HILL];
    double reac_coeff[BT_NET_MAX_TERMS], reac_rkm[BT_NET_MAX_TERMS], prod_coeff[BT_NET_MAX_TERMS];
    double r_prod, r_decay, decay_max, inv_gad, vmax, inv_keq;
};
constexpr int NET_MAX_TERMS_ALL = BT_NET_MAX_RATES * BT_NET_MAX_STOICH;
struct KNet {
    int n_species, n_rates;
    int sub_ptr[BT_NET_MAX_SPECIES + 1];
    int term_rate[NET_MAX_TERMS_ALL];
    double term_coeff[NET_MAX_TERMS_ALL];
    KNetRate rates[BT_NET_MAX_RATES];
};

// (c/Km)**n: exponents 1 and 2 take numpy's exact fast paths; 4 is (x*x)*(x*x)
// where numpy calls pow (the two differ by at most an ulp or two, and FP64 pow
// costs an order of magnitude more); other exponents call pow
__device__ __forceinline__ double hill_pow(double x, double n) {
    if (n == 1.0) return x;
    if (n == 2.0) return x * x;
    if (n == 4.0) {
        const double x2 = x * x;
        return x2 * x2;
    }
    return pow(x, n);
}

// activator term (((c/Km)**n)/(1 + (c/Km)**n)) and inhibitor term (1/(1 + (c/Km)**n))
// (get_influencers, networks.py:5184-5200, 5396-5420); the reference evaluates
// (c/Km)**n twice to the same value, here once; rkm = 1/Km
__device__ __forceinline__ double hill_act(double c, double rkm, double n) {
    const double x = hill_pow(c * rkm, n);
    return x / (1.0 + x);
}
__device__ __forceinline__ double hill_inh(double c, double rkm, double n) {
    return 1.0 / (1.0 + hill_pow(c * rkm, n));
}

// A block holds NET_CELLS cells and NET_WAVES wavefronts: lane l of every
// wavefront works on the NET_CPL cells l, l + 64, ... side by side, and
// wavefront w evaluates the rates w, w + NET_WAVES, ... from the pre-step
// concentrations (rate-uniform per wavefront, so the network description is
// read through scalar loads, once for the lane's cells, and no lane diverges
// from its wavefront). After a barrier, wavefront w updates the substances w,
// w + NET_WAVES, ...: delta[s] sums coeff * rate over the substance's terms in
// ascending rate order (np.dot(reaction_matrix, rates)), in the reference's
// order (growth/decay terms of every substance, then reactions), then
// c + delta*dt and no_negs. The concentrations and rates sit in LDS sized to
// the network at launch ((S + R) rows of NET_CELLS doubles), so a 100k-cell
// problem's blocks all fit the GPU at once.
constexpr int NET_LANES = 64;
constexpr int NET_CPL = 2;
constexpr int NET_CELLS = NET_LANES * NET_CPL;
constexpr int NET_WAVES = 4;
constexpr int NET_BLOCK = NET_LANES * NET_WAVES;

static size_t net_lds_bytes(int n_species, int n_rates) {
    return (size_t)(n_species + n_rates) * NET_CELLS * sizeof(double);
}

__global__ void __launch_bounds__(NET_BLOCK)
k_network(const KNet* __restrict__ net, double* __restrict__ cn, int n_cells, double dt,
          unsigned* flags, unsigned long long* bad_step, const long long* step_ctr, int step_ahead) {
    if (halted(flags)) return;
    extern __shared__ double net_lds[];
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    const int t = threadIdx.x % NET_LANES;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / NET_LANES));
    const int cell0 = blockIdx.x * NET_CELLS + t;
    const int S = net->n_species;
    const int R = net->n_rates;
    double* s_c = net_lds;                              // [S][NET_CELLS]
    double* s_r = net_lds + (size_t)S * NET_CELLS;      // [R][NET_CELLS]
    // cells past the end evaluate on ones and store nothing
    for (int s = w; s < S; s += NET_WAVES) {
#pragma unroll
        for (int k = 0; k < NET_CPL; ++k) {
            const int cell = cell0 + k * NET_LANES;
            s_c[s * NET_CELLS + t + k * NET_LANES] = (cell < n_cells) ? cn[(size_t)s * n_cells + cell] : 1.0;
        }
    }
    lds_barrier();
    // the concentration of substance sp in the lane's k-th cell
    auto conc = [&](int sp, int k) -> double { return s_c[sp * NET_CELLS + t + k * NET_LANES]; };
    for (int j = w; j < R; j += NET_WAVES) {
        const KNetRate& r = net->rates[j];
        // alpha = (product of activator terms) * (product of inhibitor terms), each
        // product taken left to right; an empty product is np.ones
        double act[NET_CPL], inh[NET_CPL], rate[NET_CPL];
#pragma unroll
        for (int k = 0; k < NET_CPL; ++k) act[k] = inh[k] = 1.0;
        for (int q = 0; q < r.n_act; ++q) {
            const int sp = r.act[q];
            const double rkm = r.act_rkm[q], n = r.act_n[q];
#pragma unroll
            for (int k = 0; k < NET_CPL; ++k) {
                const double h = hill_act(conc(sp, k), rkm, n);
                act[k] = (q == 0) ? h : act[k] * h;
            }
        }
        for (int q = 0; q < r.n_inh; ++q) {
            const int sp = r.inh[q];
            const double rkm = r.inh_rkm[q], n = r.inh_n[q];
#pragma unroll
            for (int k = 0; k < NET_CPL; ++k) {
                const double h = hill_inh(conc(sp, k), rkm, n);
                inh[k] = (q == 0) ? h : inh[k] * h;
            }
        }
        if (r.kind == BT_RATE_GROWTH) {
            // mod*r_prod*alpha - r_decay*cc - decay_max*1*cc, cc = c/gad_scale (networks.py:1233-1249)
#pragma unroll
            for (int k = 0; k < NET_CPL; ++k) {
                const double cc = conc(r.species, k) * r.inv_gad;
                double v = ((1.0 * r.r_prod) * act[k]) * inh[k];
                v = v - r.r_decay * cc;
                rate[k] = v - ((r.decay_max * 1.0) * 1.0) * cc;
            }
        } else {
            double x[NET_CPL];
            if (r.reversible) {
                // Q = prod(1e-3*c**k) / prod(1e-3*c**k) (mol/L), rate ~ (1.0 - Q/Keq) (networks.py:1346-1391, 1520-1524)
                double num[NET_CPL], den[NET_CPL];
                for (int q = 0; q < r.n_prod; ++q) {
                    const int sp = r.prod[q];
                    const double e = r.prod_coeff[q];
#pragma unroll
                    for (int k = 0; k < NET_CPL; ++k) {
                        const double f = 1.0e-3 * hill_pow(conc(sp, k), e);
                        num[k] = (q == 0) ? f : num[k] * f;
                    }
                }
                for (int q = 0; q < r.n_reac; ++q) {
                    const int sp = r.reac[q];
                    const double e = r.reac_coeff[q];
#pragma unroll
                    for (int k = 0; k < NET_CPL; ++k) {
                        const double f = 1.0e-3 * hill_pow(conc(sp, k), e);
                        den[k] = (q == 0) ? f : den[k] * f;
                    }
                }
#pragma unroll
                for (int k = 0; k < NET_CPL; ++k) x[k] = 1.0 - ((num[k] / den[k]) * r.inv_keq);
            } else {
                // forward coefficient: product over reactants of ((c/Km)**k)/(1 + (c/Km)**k)
                double fwd[NET_CPL];
#pragma unroll
                for (int k = 0; k < NET_CPL; ++k) fwd[k] = 1.0;
                for (int q = 0; q < r.n_reac; ++q) {
                    const int sp = r.reac[q];
                    const double rkm = r.reac_rkm[q], e = r.reac_coeff[q];
#pragma unroll
                    for (int k = 0; k < NET_CPL; ++k) {
                        const double f = hill_act(conc(sp, k), rkm, e);
                        fwd[k] = (q == 0) ? f : fwd[k] * f;
                    }
                }
#pragma unroll
                for (int k = 0; k < NET_CPL; ++k) x[k] = fwd[k] - (0.0 * 0.0);
            }
#pragma unroll
            for (int k = 0; k < NET_CPL; ++k) rate[k] = ((r.vmax * act[k]) * inh[k]) * x[k];
        }
#pragma unroll
        for (int k = 0; k < NET_CPL; ++k) s_r[j * NET_CELLS + t + k * NET_LANES] = rate[k];
    }
    lds_barrier();
    bool bad = false;
    for (int s = w; s < S; s += NET_WAVES) {
        double d[NET_CPL];
#pragma unroll
        for (int k = 0; k < NET_CPL; ++k) d[k] = 0.0;
        for (int q = net->sub_ptr[s]; q < net->sub_ptr[s + 1]; ++q) {
            const double co = net->term_coeff[q];
            const int jr = net->term_rate[q];
#pragma unroll
            for (int k = 0; k < NET_CPL; ++k) d[k] = d[k] + co * s_r[jr * NET_CELLS + t + k * NET_LANES];
        }
#pragma unroll
        for (int k = 0; k < NET_CPL; ++k) {
            const int cell = cell0 + k * NET_LANES;
            double v = conc(s, k) + d[k] * dt;
            if (v < 0.0) v = 0.0;                    // stb.no_negs clamps negatives ...
            if (cell < n_cells) {
                if (isnan(v)) bad = true;            // ... and raises on NaN
                cn[(size_t)s * n_cells + cell] = v;
            }
        }
    }
    if (bad) s_bad = 1;
    lds_barrier();
    if (threadIdx.x == 0 && s_bad) {
        // the step's index: step_ctr counts the steps whose exchange has run
        atomicOr(flags, 1u);
        atomicMin(bad_step, (unsigned long long)(*step_ctr - step_ahead));
    }
}

// ---------------------------------------------------------------------------
// sampled-step observables
// ---------------------------------------------------------------------------
struct ObsMemArgs {
    const int32_t* mem_ptr; const int32_t* nn_mem; const double* mem_sa; const double* mem_nx;
    const double* mem_ny; const double* cell_sa; const double* vm_cell; const double* vm_gj;
    const int32_t* mem_cell; const int32_t* tile_ptr;
    const double* fm; const double* fgj; const double* fch;
    double* Jn; double* I_mem; double* E_gj_x; double* E_gj_y;
    double* J_cell_x; double* J_cell_y; double* vm_ave;
    double smooth_cells;
    // applied boundary voltage per square (NULL: none) and map_mem2ecm: the
    // membranes' Vmem is their cell's less Phi_b at their square
    const double* Phi_b; const double* Phi_b_gj; const int32_t* mem_sq;
    double* vm_mem;           // Vmem per membrane (write2storage's vm), or NULL
};

// raw normal current of membrane m: -sum(zF*f_mem) + extra_J_mem (channels) +
// sum(zF*f_gj), in the reference's order (ion_current.py:18-30)
__device__ __forceinline__ double obs_jn_raw(const ObsMemArgs& A, const KParams& P, int m) {
    const int M = P.n_mems;
    double extra = 0.0;                       // extra_J_mem: channel currents
    for (int ch = 0; ch < P.n_chan; ++ch) {
        const int ion = P.chan_ion[ch];
        extra = extra + ((-A.fch[(size_t)ch * M + m]) * P.F) * P.zq[ion];
    }
    double d = P.zF[0] * A.fm[m];
    double g = P.zF[0] * A.fgj[m];
    for (int i = 1; i < P.n_ions; ++i) {
        d = d + P.zF[i] * A.fm[(size_t)i * M + m];
        g = g + P.zF[i] * A.fgj[(size_t)i * M + m];
    }
    return (-d + extra) + g;
}

// get_current cell/membrane part (ion_current.py:18-47), E_gj (sim.py:2166-2172,
// from the Vmem the step started with, vm_gj), vm_ave (sim.py:2038): one
// wavefront per tile of whole cells, as k_mem_cell. Membrane lanes evaluate,
// cell lanes take the sums over their membranes in membrane order from LDS.
__global__ void __launch_bounds__(TMB)
k_obs_mem(ObsMemArgs A, KParams P) {
    __shared__ double s_j[TMB], s_sa[TMB], s_jx[TMB], s_jy[TMB];
    __shared__ double s_ave[TMB], s_wm[TMB], s_wo[TMB];      // per cell of the tile
    __shared__ double s_vm[TMB];                              // Vmem of each membrane
    __shared__ int s_lc[TMB];                                 // cell of each membrane, in the tile
    const int tile = blockIdx.x;
    const int c0 = A.tile_ptr[tile], c1 = A.tile_ptr[tile + 1];
    const int m0 = A.mem_ptr[c0], m1 = A.mem_ptr[c1];
    const int j = threadIdx.x;
    const int m = m0 + j, c = c0 + j;
    int ca = 0, cb = 0;
    if (c < c1) {
        ca = A.mem_ptr[c] - m0;
        cb = A.mem_ptr[c + 1] - m0;
    }
    if (m < m1) {
        s_j[j] = obs_jn_raw(A, P, m);
        s_sa[j] = A.mem_sa[m];
        s_lc[j] = A.mem_cell[m] - c0;
    }
    lds_barrier();
    if (c < c1) {
        double s = 0.0;
        for (int q = ca; q < cb; ++q) s = (q == ca) ? s_j[q] * s_sa[q] : s + s_j[q] * s_sa[q];
        s_ave[j] = s / A.cell_sa[c];
        const double nm = A.smooth_cells * (double)(cb - ca);
        s_wm[j] = (nm - 1) / nm;
        s_wo[j] = 1 / nm;
    }
    lds_barrier();
    if (m < m1) {
        const int lc = s_lc[j];
        const double jn = s_wm[lc] * s_j[j] + s_ave[lc] * s_wo[lc];
        A.Jn[m] = jn;
        A.I_mem[m] = -jn * s_sa[j];
        s_jx[j] = (jn * A.mem_nx[m]) * s_sa[j];
        s_jy[j] = (jn * A.mem_ny[m]) * s_sa[j];
        double vgj;
        if (A.Phi_b) {
            const int nn = A.nn_mem[m];
            vgj = (A.vm_gj[A.mem_cell[nn]] - A.Phi_b_gj[A.mem_sq[nn]]) -
                  (A.vm_gj[c0 + lc] - A.Phi_b_gj[A.mem_sq[m]]);
            s_vm[j] = A.vm_cell[c0 + lc] - A.Phi_b[A.mem_sq[m]];
        } else {
            vgj = A.vm_gj[A.mem_cell[A.nn_mem[m]]] - A.vm_gj[c0 + lc];
            s_vm[j] = A.vm_cell[c0 + lc];
        }
        const double egj = -vgj / P.gj_len;
        if (A.vm_mem) A.vm_mem[m] = s_vm[j];
        A.E_gj_x[m] = egj * A.mem_nx[m];
        A.E_gj_y[m] = egj * A.mem_ny[m];
    }
    lds_barrier();
    if (c < c1) {
        double sx = s_jx[ca], sy = s_jy[ca], svm = s_vm[ca];
        for (int q = ca + 1; q < cb; ++q) {
            sx = sx + s_jx[q];
            sy = sy + s_jy[q];
            svm = svm + s_vm[q];
        }
        A.J_cell_x[c] = sx / A.cell_sa[c];
        A.J_cell_y[c] = sy / A.cell_sa[c];
        A.vm_ave[c] = svm / (double)(cb - ca);
    }
}

// env current J_env_o = F*sum(z*fluxes_env) (ion_current.py:53-58) and the two
// right-hand sides of HH_Decomp (sim_toolbox.py:1253-1282) on the interior,
// zero-padded to the GEMM tile (boundary values are zero without bound_V)
__global__ void __launch_bounds__(BLOCK)
k_obs_env_rhs(const double* fex, const double* fey, double* jxo, double* jyo, KParams P) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n_sq) return;
    double x = (P.F * P.zq[0]) * fex[k], y = (P.F * P.zq[0]) * fey[k];
    for (int i = 1; i < P.n_ions; ++i) {
        x = x + (P.F * P.zq[i]) * fex[(size_t)i * P.n_sq + k];
        y = y + (P.F * P.zq[i]) * fey[(size_t)i * P.n_sq + k];
    }
    jxo[k] = x;
    jyo[k] = y;
}

// fd.divergence of (Fx, Fy) at interior square (r, c) (finitediff.py:1268-1311)
__device__ __forceinline__ double div_at(const double* Fx, const double* Fy, int r, int c, int ny, int nx,
                                         double dl, double d2, double sx, double sy) {
    auto X = [&](int rr, int cc) { return sx * Fx[rr * nx + cc]; };
    auto Y = [&](int rr, int cc) { return sy * Fy[rr * nx + cc]; };
    double gx, gy;
    if (c == 0) gx = (X(r, 0) - X(r, 1)) / dl;
    else if (c == nx - 1) gx = (X(r, nx - 2) - X(r, nx - 1)) / dl;
    else gx = -(X(r, c - 1) - X(r, c + 1)) / d2;
    if (r == 0) gy = -(Y(1, c) - Y(0, c)) / dl;
    else if (r == ny - 1) gy = -(Y(ny - 1, c) - Y(ny - 2, c)) / dl;
    else gy = -(Y(r - 1, c) - Y(r + 1, c)) / d2;
    return gx + gy;
}

// One cluster's share of the batched Helmholtz-Hodge solves: its interior
// (n1 x n2 squares), padded to the GEMM tile (n1p x n2p), and the offset of
// its padded matrices in the batch's workspace (in units of elements; each
// workspace buffer holds, per cluster, the stream-function matrix and then the
// potential matrix, pad = n1p * n2p elements each).
struct ObsQ { int n1, n2, n1p, n2p; long long off; };

// the cluster whose workspace range holds offset w: the last q with
// off[q] <= w (binary search over the batch's ascending offsets)
__device__ __forceinline__ int obs_cluster(const ObsQ* Q, int nq, long long w) {
    int lo = 0, hi = nq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (Q[mid].off <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// rhs of the stream-function solve (-div(-Jy, Jx)) and of the potential solve
// (div(Jx, Jy)) of every cluster's interior, written row-major into its two
// padded matrices with zeros in the padding: one thread per padded interior
// element e of the batch, n_pad = sum of n1p * n2p over the clusters (the
// matrices need no clearing beforehand)
__global__ void __launch_bounds__(BLOCK)
k_obs_rhs(const double* jxo, const double* jyo, double* Bw, const ObsQ* Q, int nq, long long n_pad,
          const KCluster* kcl) {
    const long long e = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n_pad) return;
    // a cluster's two matrices take 2 n1p n2p elements, so element e of the
    // interiors lies in the cluster whose range holds offset 2 e
    const int q = obs_cluster(Q, nq, 2 * e);
    const ObsQ g = Q[q];
    const int pad = g.n1p * g.n2p;
    const int loc = (int)(e - g.off / 2);          // element of the cluster's padded interior
    const int rr = loc / g.n2p, cc = loc % g.n2p;
    double b1 = 0.0, b2 = 0.0;
    if (rr < g.n1 && cc < g.n2) {
        const KCluster& K = kcl[q];
        const double* jx = jxo + K.sq_off;
        const double* jy = jyo + K.sq_off;
        // Jxr = -Jy, Jyr = Jx
        const double dr = div_at(jy, jx, rr + 1, cc + 1, K.ny, K.nx, K.delta, K.two_delta, -1.0, 1.0);
        const double dd = div_at(jx, jy, rr + 1, cc + 1, K.ny, K.nx, K.delta, K.two_delta, 1.0, 1.0);
        b1 = -dr;
        b2 = dd;
    }
    Bw[g.off + loc] = b1;
    Bw[g.off + pad + loc] = b2;
}

// gradients of the stream function AA and potential BB -> HH_Decomp outputs,
// over every square of the batch. AA and BB are the interior solutions
// (padded, ld = n2p) with zero boundary rows and columns (their boundary
// right-hand sides are zero without an applied voltage).
__global__ void __launch_bounds__(BLOCK)
k_obs_hh_out(const double* Xw, const ObsQ* Q, const int32_t* sq_cluster, const KCluster* kcl, double* Jx,
             double* Jy, double* Jtx, double* Jty, double* Bf, double mu, int n_sq, const double* Phi_b) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n_sq) return;
    const int q = sq_cluster[k];
    const KCluster& K = kcl[q];
    const ObsQ g = Q[q];
    const double* AA = Xw + g.off;
    const double* BB = AA + g.n1p * g.n2p;
    const int ny = K.ny, nx = K.nx, loc = k - K.sq_off, r = loc / nx, c = loc % nx;
    auto F = [&](const double* X, int rr, int cc) {
        return (rr == 0 || cc == 0 || rr == ny - 1 || cc == nx - 1) ? 0.0 : X[(rr - 1) * g.n2p + (cc - 1)];
    };
    // the potential with an applied boundary voltage: the boundary rows of its
    // solve hold -bound_V, so it is the zero-boundary solution less Phi_b (the
    // harmonic extension of bound_V with the same corner rule)
    const double* Pb = (Phi_b && K.phib) ? Phi_b + K.sq_off : nullptr;
    auto FB = [&](int rr, int cc) { return Pb ? F(BB, rr, cc) - Pb[rr * nx + cc] : F(BB, rr, cc); };
    auto grad = [&](auto G, double& gx, double& gy) {
        if (c == 0) gx = (G(r, 1) - G(r, 0)) / K.delta;
        else if (c == nx - 1) gx = (G(r, nx - 1) - G(r, nx - 2)) / K.delta;
        else gx = -(G(r, c - 1) - G(r, c + 1)) / K.two_delta;
        if (r == 0) gy = (G(1, c) - G(0, c)) / K.delta;
        else if (r == ny - 1) gy = (G(ny - 1, c) - G(ny - 2, c)) / K.delta;
        else gy = -(G(r - 1, c) - G(r + 1, c)) / K.two_delta;
    };
    double gAx, gAy, Gx, Gy;
    grad([&](int rr, int cc) { return F(AA, rr, cc); }, gAx, gAy);
    grad(FB, Gx, Gy);
    const double Fx = -gAy, Fy = gAx;
    Jx[k] = Fx;
    Jy[k] = Fy;
    Jtx[k] = Fx + Gx;
    Jty[k] = Fy + Gy;
    Bf[k] = F(AA, r, c) * mu;
}

// One 16x16 tile of C = A*B (row-major, all sizes multiples of 16), times S
// element by element when S is given, with v_mfma_f64_16x16x4f64: a block of
// four wavefronts, wavefront w taking the quarter [w K/4, (w+1) K/4) of the
// inner dimension (a multiple of 4), its operands loaded sixteen k-steps at a
// time before their MFMAs; the four partial tiles are then added in wavefront
// order through LDS (a fixed order: the result does not vary between runs).
// Lane l holds A[l%16][k0+l/16] and B[k0+l/16][l%16]; result element r of
// lane l is C[l/16+4*r][l%16].
typedef double v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void gemm_tile(const double* A, const double* B, double* C, const double* S, int n,
                                          int kd, int tile, double (*part)[64][4]) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int tiles_n = n / 16;
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    const int row = tm * 16 + (lane % 16), col = tn * 16 + (lane % 16), kk = lane / 16;
    const int kq = kd / 4;
    const int ke = (wave + 1) * kq;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    int k0 = wave * kq;
    for (; k0 + 64 <= ke; k0 += 64) {
        double a[16], b[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            a[u] = A[(size_t)row * kd + k0 + 4 * u + kk];
            b[u] = B[(size_t)(k0 + 4 * u + kk) * n + col];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
    for (; k0 < ke; k0 += 4) {
        const double a = A[(size_t)row * kd + k0 + kk];
        const double b = B[(size_t)(k0 + kk) * n + col];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) part[wave - 1][lane][r] = acc[r];
    }
    lds_barrier();
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double sum = acc[r];
            sum = sum + part[0][lane][r];
            sum = sum + part[1][lane][r];
            sum = sum + part[2][lane][r];
            const size_t o = (size_t)(tm * 16 + kk + 4 * r) * n + col;
            C[o] = S ? sum * S[o] : sum;
        }
    }
}

__global__ void __launch_bounds__(256)
k_gemm_f64(const double* A, const double* B, double* C, int m, int n, int kd) {
    __shared__ double part[3][64][4];            // partial tiles of wavefronts 1..3
    const int tile = blockIdx.x;                 // block-uniform
    if (tile >= (m / 16) * (n / 16)) return;
    gemm_tile(A, B, C, nullptr, n, kd, tile, part);
}

// Many independent products in one launch (the clusters' DST stages): block b
// computes tile b - tile0 of product block_prob[b].
struct GemmProb { const double* A; const double* B; double* C; const double* S; int m, n, k, tile0; };

__global__ void __launch_bounds__(256)
k_gemm_grouped(const GemmProb* probs, const int32_t* block_prob) {
    __shared__ double part[3][64][4];
    const GemmProb P = probs[block_prob[blockIdx.x]];
    gemm_tile(P.A, P.B, P.C, P.S, P.n, P.k, (int)blockIdx.x - P.tile0, part);
}

// per-cluster max over cells and ions of |c - c_prev| / |c| (init residual)
__global__ void __launch_bounds__(BLOCK)
k_residual(const double* c, const double* c_prev, const int32_t* cell_cluster, unsigned long long* rmax,
           int n_cells, int n_ions) {
    const int cidx = blockIdx.x * BLOCK + threadIdx.x;
    if (cidx >= n_cells) return;
    double r = 0.0;
    for (int i = 0; i < n_ions; ++i) {
        const double x = c[(size_t)i * n_cells + cidx];
        const double d = fabs(x - c_prev[(size_t)i * n_cells + cidx]);
        const double q = (x != 0.0) ? d / fabs(x) : d;
        if (q > r || isnan(q)) r = isnan(q) ? INFINITY : q;
    }
    // non-negative doubles order like their bit patterns
    atomicMax(&rmax[cell_cluster[cidx]], (unsigned long long)__double_as_longlong(r));
}

// ---------------------------------------------------------------------------
// k_step_persistent: many steps in one cooperative launch
// ---------------------------------------------------------------------------
// Each step's three dependent phases (env field + diffusion, membranes and
// cells, exchange) run in one launch whose blocks all stay resident
// (hipLaunchCooperativeKernel) and meet at a grid barrier between phases: the
// per-launch drain and ramp of the three-kernel step are gone, and a launch
// runs any even number of steps. Work is dealt out in grid-stride order: env
// (tile, ion group) items per block, cell tiles per wavefront (each wavefront
// owns one tile at a time, with the membrane kernel's single-wavefront form),
// squares with membranes per thread.
constexpr int PB = 256;                      // threads per persistent block
constexpr int PW = PB / TMB;                 // wavefronts per persistent block
#ifndef PERSIST_WAVES_PER_SIMD
#define PERSIST_WAVES_PER_SIMD 4             // occupancy the register budget is held to
#endif
constexpr unsigned BT_FLAG_BARRIER = 2u;     // flags bit: a grid barrier gave up waiting
constexpr unsigned BARRIER_SPINS = 1u << 24; // about a second of s_sleep before it gives up

struct PersistArgs {
    FieldArgs field[2];                      // by step parity
    MemCellArgs mem[2];
    ExArgs ex[2];
    unsigned* bar;                           // [0] arrivals, [1] generation
    int n_env_tiles, n_groups, n_tiles, n_steps;
};

// LDS of one membrane-phase wavefront, in doubles: flux*sa and GJ flux per ion,
// channel flux per channel, and the tile's NaN flag
template <int NI>
constexpr int mem_wave_lds() { return (2 * NI + BT_MAX_CHAN) * TMB + 1; }

// Grid barrier of a cooperative launch. Thread 0 of each block reads the
// generation, arrives, and the last arrival resets the count and opens the next
// generation; the others wait for it. The agent-scope fences make every block's
// global writes before the barrier visible to every block after it (across the
// XCDs' L2 caches). A wait that exceeds BARRIER_SPINS sets BT_FLAG_BARRIER and
// the block goes on, so every wave reaches the end of the kernel whatever
// happens. Returns false once the flag is set.
__device__ __forceinline__ bool grid_sync(unsigned* bar, unsigned* flags) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __threadfence();
        const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (atomicAdd(&bar[0], 1u) == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            atomicAdd(&bar[1], 1u);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > BARRIER_SPINS) {
                    atomicOr(flags, BT_FLAG_BARRIER);
                    break;
                }
            }
        }
        __threadfence();
        if (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & BT_FLAG_BARRIER) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <int NI, typename FT, bool RATES>
__global__ void __launch_bounds__(PB, PERSIST_WAVES_PER_SIMD)
k_step_persistent(PersistArgs PA, KParams P) {
    constexpr int MW = mem_wave_lds<NI>();
    constexpr int LDS = ENV_LDS > PW * MW ? ENV_LDS : PW * MW;
    __shared__ double lds[LDS];
    __shared__ int s_stop;
    const int nb = gridDim.x;
    const int wave = threadIdx.x / TMB;
    const int gtid = blockIdx.x * PB + threadIdx.x;
    const int nthr = nb * PB;
    const int n_env_items = PA.n_env_tiles * PA.n_groups;
    // this wavefront's slice of the LDS in the membrane phase
    double* wl = lds + wave * MW;
    const MemLds L{reinterpret_cast<double (*)[TMB]>(wl), reinterpret_cast<double (*)[TMB]>(wl + NI * TMB),
                   reinterpret_cast<double (*)[TMB]>(wl + 2 * NI * TMB), reinterpret_cast<double (*)[TMB]>(wl),
                   reinterpret_cast<int*>(wl + (2 * NI + BT_MAX_CHAN) * TMB)};
    // one step of parity PAR (a compile-time index into the argument sets);
    // false once a grid barrier gave up
    auto step = [&](auto par_c) -> bool {
        constexpr int par = decltype(par_c)::value;
        const FieldArgs& FA = PA.field[par];
        // phase A: env field and diffusion (and the gating coefficients). The
        // flags change only in the membrane phase, so one read decides for the block.
        if (threadIdx.x == 0) s_stop = halted(FA.flags) ? 1 : 0;
        lds_barrier();
        if (!s_stop) {
            for (int it = blockIdx.x; it < n_env_items; it += nb) {
                const int item = (nb % N_XCD == 0) ? xcd_chunk(it, n_env_items) : it;
                env_field_tile<true, false>(FA, P, item % PA.n_env_tiles, item / PA.n_env_tiles, PA.n_groups, lds,
                                            -1, 0, false);
                lds_barrier();
            }
            if (RATES) {
                const int items = P.n_chan * P.n_cells;
                for (int t_it = gtid; t_it < items; t_it += nthr) gate_coef_item(FA, P, t_it, FA.vm[t_it % P.n_cells]);
            }
        }
        if (!grid_sync(PA.bar, PA.mem[par].flags)) return false;
        // phase B: membranes and cells, one tile per wavefront at a time; the
        // halt check is uniform over the wavefront
        {
            const MemCellArgs& MA = PA.mem[par];
            const bool stop = __builtin_amdgcn_readfirstlane(halted(MA.flags) ? 1 : 0) != 0;
            if (!stop) {
                const int nw = nb * PW;
                for (int w = blockIdx.x * PW + wave; w < PA.n_tiles; w += nw) {
                    mem_cell_wave<NI, FT, 1, 0, false, true>(MA, P, w, L);
                    tile_sync<true>();
                }
            }
        }
        if (!grid_sync(PA.bar, PA.mem[par].flags)) return false;
        // phase C: the exchange of the squares with membranes; the step counter
        {
            const ExArgs& XA = PA.ex[par];
            if (!halted(XA.flags)) {
                if (gtid == 0) advance_step(XA);
                for (int q = gtid; q < XA.n_live; q += nthr) exchange_square(XA, P, q);
            }
        }
        return grid_sync(PA.bar, PA.mem[par].flags);
    };
    for (int s = 0; s < PA.n_steps; s += 2) {
        if (!step(std::integral_constant<int, 0>{})) break;
        if (!step(std::integral_constant<int, 1>{})) break;
    }
}

}  // namespace

// ===========================================================================
// context
// ===========================================================================
struct bt_ctx {
    bt_params prm;
    bt_buffers buf;
    KParams kp;
    std::vector<bt_cluster> clusters;
    std::vector<KCluster> kcl_host;
    KCluster* kcl = nullptr;        // device copy
    int* env_tiles = nullptr;       // device (cluster, r0, c0) per env tile
    int n_env_tiles = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int n_tiles = 0;
    // workspace
    double* c_cell_b = nullptr; double* c_lag_b = nullptr; double* vm_b = nullptr;
    double* c_mid = nullptr; void* fsa = nullptr; double* fch = nullptr; double* u = nullptr;
    double* u_b = nullptr;    // surface charge of odd steps with the fused exchange
    double* rates = nullptr;  // [n_chan][4][n_cells] gating rates of the current step
    double* c_env_b = nullptr;           // env state of odd steps
    int32_t* sq_live = nullptr;          // squares the exchange updates, ascending
    uint8_t* live_flag = nullptr;        // per square: 1 if in sq_live
    int n_live = 0;
    unsigned* flags = nullptr; unsigned long long* bad_step = nullptr; long long* step_ctr = nullptr;
    // step graphs per stream group g: one pair of steps, and tune.graph_pairs pairs
    hipGraph_t graph[BT_MAX_STREAMS] = {}; hipGraphExec_t gexec[BT_MAX_STREAMS] = {};
    hipGraph_t graph_n[BT_MAX_STREAMS] = {}; hipGraphExec_t gexec_n[BT_MAX_STREAMS] = {};
    // stream groups (tune.streams > 1): per cluster q, the first of its cell
    // tiles, env tiles and live squares and of its cells (n_clusters + 1
    // entries each; empty where the clusters' work lists are not in cluster
    // order); the groups' streams ([0] is the caller's) and fork/join events
    std::vector<int> cl_tile0, cl_env0, cl_live0;
    std::vector<int> cl_mem_lo, cl_mem_hi;   // each cluster's membrane range
    bool any_phib = false;                   // some cluster has an applied boundary voltage
    std::vector<long long> cl_cell0;
    hipStream_t gstream[BT_MAX_STREAMS] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[BT_MAX_STREAMS] = {};
    bt_tuning tune{};                  // launch variants of the step kernels
    bool autotuned = false;
    unsigned* bar = nullptr;           // grid barrier of the persistent step kernel
    int pers_grid = -1;                // its cooperative grid (-1 not yet known, 0 unavailable)
    long long steps_done = 0;
    const bt_obs* capture = nullptr;   // set while enqueueing an observed step
    // observables workspace (allocated on first use)
    std::vector<void*> obs_alloc;        // every observables buffer, freed in bt_destroy
    bool obs_ready = false;
    double* jxo = nullptr; double* jyo = nullptr;
    // the batched Poisson solves: per-cluster padded matrices (right-hand
    // sides, two products' workspace), the cluster table, and the four DST
    // stages as grouped products (n_prob products each, n_gemm_blocks tiles)
    double* Bw = nullptr; double* T1w = nullptr; double* T2w = nullptr;
    ObsQ* obs_q = nullptr; long long n_pad = 0;
    GemmProb* gemm_probs = nullptr; int32_t* block_prob = nullptr; int n_prob = 0; int n_gemm_blocks = 0;
    // Vmem per cell before the last observed step: update_gj's E_gj uses it (sim.py:2166-2172)
    double* vm_prev = nullptr; bool vm_prev_ok = false;
    // applied boundary voltage: the Phi_b staged by bt_stage_boundary for the
    // end of the next step, and the Phi_b in effect during the last observed
    // step (its E_gj takes Vmem from before the step)
    double* phib_stage = nullptr; bool phib_staged = false;
    double* phib_prev = nullptr;
    // reaction network (device copy of the description) and its state buffer
    KNet* net = nullptr; double* c_net = nullptr; int net_species = 0, net_rates = 0;
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
    k.n_cells = p.n_cells; k.n_mems = p.n_mems; k.n_sq = p.n_squares; k.n_ions = p.n_ions;
    k.i_na = p.i_na; k.i_k = p.i_k; k.i_ca = p.i_ca;
    k.n_chan = p.n_chan;
    for (int i = 0; i < BT_MAX_IONS; ++i) k.ion_chan[i] = -1;
    for (int c = 0; c < p.n_chan; ++c) {
        k.chan_kind[c] = p.chan_kind[c];
        k.chan_ion[c] = p.chan_ion[c];
        k.chan_max_dm[c] = p.chan_max_dm[c];
        k.ion_chan[p.chan_ion[c]] = c;
    }
    k.chan_dt = p.chan_dt;
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
    k.f_rt = p.F / (p.R * p.T);
    k.f_rtgj = p.F / (p.R * p.T_gj);
    k.inv_tm = 1.0 / p.tm;
    k.inv_gjlen = 1.0 / p.gj_len;
    k.inv_rt = 1.0 / (p.R * p.T);
    k.inv_kmna = 1.0 / p.KmNK_Na;
    k.inv_kmk = 1.0 / p.KmNK_K;
    k.inv_kmca = 1.0 / p.KmCa_Ca;
    k.f_rt32 = (float)(p.F / (p.R * p.T));
    k.f_rtgj32 = (float)(p.F / (p.R * p.T_gj));
    k.inv_tm32 = (float)(1.0 / p.tm);
    k.inv_gjlen32 = (float)(1.0 / p.gj_len);
    k.inv_rt32 = (float)(1.0 / (p.R * p.T));
    k.inv_kmna32 = (float)(1.0 / p.KmNK_Na);
    k.inv_kmk32 = (float)(1.0 / p.KmNK_K);
    k.inv_kmca32 = (float)(1.0 / p.KmCa_Ca);
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
    for (int j = 0; j < 5; ++j) k.gw[j] = p.gauss_w[j];
}

static KCluster derive_cluster(const bt_params& p, const bt_cluster& b) {
    KCluster k;
    memset(&k, 0, sizeof(k));
    k.ny = b.grid_ny; k.nx = b.grid_nx; k.sq_off = b.sq_off;
    k.delta = b.delta; k.two_delta = 2.0 * b.delta;
    k.env_vol = p.cell_height * (b.delta * b.delta);
    k.surf_scale = p.cell_height * (b.delta * b.delta);
    k.mean_sa = b.mean_sa;
    k.kee = (b.ko_env * p.eo) * p.er;
    k.screen = (2.0 / (b.ko_env * b.delta)) * (p.cell_radius / p.true_cell_size);
    k.nak_coef = -3.0 * p.NaK_block * b.alpha_NaK;
    for (int i = 0; i < BT_MAX_IONS; ++i) k.dm_u[i] = k.denv_u[i] = NAN;
    k.gjo_u = NAN;
    return k;
}

// Default wavefronts per tile of k_mem_cell: two (ions split between them)
// while one wavefront per tile would leave the SIMDs with few waves to switch
// between; pairs of steps per replayed graph (fewer graph launches).
constexpr int BT_MEM_WG_TILES = 4096;
constexpr int BT_GRAPH_PAIRS = 8;
constexpr int BT_MAX_GRAPH_PAIRS = 64;
constexpr int BT_PERSIST_STEPS = 512;   // steps per cooperative launch of the persistent kernel (even)
constexpr int BT_AUTOTUNE_TRIALS = 13;

// A stream group's share of the step's work lists: its cell tiles [t0, t0 +
// nt), env tiles [e0, e0 + ne) and live squares [l0, l0 + nl).
struct Span { int t0, nt, e0, ne, l0, nl; };

// The span of group g of G (G = 1: the whole problem). Group g takes the
// clusters [q_g, q_{g+1}), q_g the first cluster whose cells start at or after
// g/G of all cells, each group keeping at least one cluster.
static Span group_span(const bt_ctx* c, int G, int g) {
    if (G <= 1) return Span{0, c->n_tiles, 0, c->n_env_tiles, 0, c->n_live};
    const int nq = (int)c->clusters.size();      // at least G (streams_allowed)
    // b[gg]: the first cluster of group gg, strictly increasing, with room
    // for one cluster in each later group
    int b[BT_MAX_STREAMS + 1];
    b[0] = 0;
    for (int gg = 1; gg < G; ++gg) {
        const long long target = c->cl_cell0[nq] * gg / G;
        const int q = (int)(std::lower_bound(c->cl_cell0.begin(), c->cl_cell0.end(), target) - c->cl_cell0.begin());
        b[gg] = std::max(std::min(q, nq - (G - gg)), b[gg - 1] + 1);
    }
    b[G] = nq;
    const int qa = b[g], qz = b[g + 1];
    return Span{c->cl_tile0[qa], c->cl_tile0[qz] - c->cl_tile0[qa], c->cl_env0[qa], c->cl_env0[qz] - c->cl_env0[qa],
                c->cl_live0[qa], c->cl_live0[qz] - c->cl_live0[qa]};
}

template <int NI, int WG, bool OBS>
static void launch_mem_cell_o(bt_ctx* c, const MemCellArgs& a, hipStream_t st, int nt) {
    const size_t lds = (size_t)c->prm.n_chan * TMB * sizeof(double);     // s_ch
    if (c->prm.flux_mode == BT_FLUX_FP32)
        hipLaunchKernelGGL((k_mem_cell<NI, float, WG, OBS>), dim3(nt), dim3(TMB * WG), lds, st, a, c->kp);
    else
        hipLaunchKernelGGL((k_mem_cell<NI, double, WG, OBS>), dim3(nt), dim3(TMB * WG), lds, st, a, c->kp);
}
template <int NI, int WG>
static void launch_mem_cell_w(bt_ctx* c, const MemCellArgs& a, hipStream_t st, int nt) {
    if (a.obs_fm) launch_mem_cell_o<NI, WG, true>(c, a, st, nt);
    else launch_mem_cell_o<NI, WG, false>(c, a, st, nt);
}
template <int NI>
static void launch_mem_cell_t(bt_ctx* c, const MemCellArgs& a, hipStream_t st, int nt) {
    if (c->tune.mem_waves == 2) launch_mem_cell_w<NI, 2>(c, a, st, nt);
    else launch_mem_cell_w<NI, 1>(c, a, st, nt);
}

// k_mem_cell over nt cell tiles (a.tile_ptr and a.tile_cluster start at the first)
static void launch_mem_cell(bt_ctx* c, const MemCellArgs& a, hipStream_t st, int nt) {
    switch (c->prm.n_ions) {
        case 1: launch_mem_cell_t<1>(c, a, st, nt); break;
        case 2: launch_mem_cell_t<2>(c, a, st, nt); break;
        case 3: launch_mem_cell_t<3>(c, a, st, nt); break;
        case 4: launch_mem_cell_t<4>(c, a, st, nt); break;
        case 5: launch_mem_cell_t<5>(c, a, st, nt); break;
        case 6: launch_mem_cell_t<6>(c, a, st, nt); break;
        case 7: launch_mem_cell_t<7>(c, a, st, nt); break;
        default: launch_mem_cell_t<8>(c, a, st, nt); break;
    }
}

static inline int grid_blocks(int n) { return (n + BLOCK - 1) / BLOCK; }

// diffuse: the step's env field and diffusion, and the gating rates from vm;
// else the field alone (v_env, E) from u
static FieldArgs field_args(bt_ctx* c, bool diffuse, const bt_obs* ob, const double* vm, const double* env_in,
                            double* env_out) {
    const bt_buffers& B = c->buf;
    return FieldArgs{c->env_tiles, c->kcl, c->u, env_in ? env_in : B.c_env, B.D_env, c->c_mid, B.v_env, B.E_x, B.E_y,
                     c->flags, ob ? ob->fluxes_env_x : nullptr, ob ? ob->fluxes_env_y : nullptr,
                     vm, (diffuse && c->prm.n_chan > 0) ? c->rates : nullptr, c->live_flag, env_out, B.Phi_b};
}

// k_env_field over ne env tiles (f.env_tiles starts at the first; ne = 0: all)
static void launch_field_args(bt_ctx* c, bool diffuse, const FieldArgs& f, hipStream_t st, int ne) {
    // diffusion: one block per (tile, env_ions_per_block ions): few ions per
    // block while tiles alone would leave CUs idle, else one block per tile
    // looping over the ions; field only: one block per tile
    const int ipb = c->tune.env_ions_per_block;
    dim3 grid(ne > 0 ? ne : c->n_env_tiles, diffuse ? (c->prm.n_ions + ipb - 1) / ipb : 1);
    if (diffuse && f.u_out)
        hipLaunchKernelGGL((k_env_field<true, false, true>), grid, dim3(TY * TX), 0, st, f, c->kp);
    else if (diffuse && f.rates)
        hipLaunchKernelGGL((k_env_field<true, true>), grid, dim3(TY * TX), 0, st, f, c->kp);
    else if (diffuse)
        hipLaunchKernelGGL((k_env_field<true, false>), grid, dim3(TY * TX), 0, st, f, c->kp);
    else
        hipLaunchKernelGGL((k_env_field<false, false>), grid, dim3(TY * TX), 0, st, f, c->kp);
}

static void launch_field(bt_ctx* c, bool diffuse, const bt_obs* ob = nullptr, const double* vm = nullptr,
                         const double* env_in = nullptr, double* env_out = nullptr) {
    launch_field_args(c, diffuse, field_args(c, diffuse, ob, vm, env_in, env_out), c->stream, 0);
}

// The kernel arguments of one step. parity 0: cell state caller -> workspace;
// parity 1: workspace -> caller (the env state ping-pongs likewise).
struct StepArgs {
    FieldArgs field;
    MemCellArgs mem;
    ExArgs ex;
};

// G, g: stream group g of G (G = 1: the whole problem); fused: the env kernel
// does the exchange (launched by enqueue_step with that tuning; the persistent
// kernel keeps its own exchange phase)
static StepArgs step_args(bt_ctx* c, int parity, const bt_obs* ob, int count_step, int G = 1, int g = 0,
                          bool fused = false) {
    const bt_buffers& B = c->buf;
    double* ci = parity ? c->c_cell_b : B.c_cell;
    double* li = parity ? c->c_lag_b : B.c_lag;
    double* vi = parity ? c->vm_b : B.vm_cell;
    double* co = parity ? B.c_cell : c->c_cell_b;
    double* lo = parity ? B.c_lag : c->c_lag_b;
    double* vo = parity ? B.vm_cell : c->vm_b;
    double* ei = parity ? c->c_env_b : B.c_env;
    double* eo = parity ? B.c_env : c->c_env_b;
    StepArgs a{field_args(c, true, ob, vi, ei, eo),
               MemCellArgs{B.mem_ptr, B.mem_cell, B.nn_mem, B.nn_cell, B.mem_sq, B.mem_pos, B.tile_ptr,
                           B.tile_cluster, c->kcl,
                           B.mem_sa, B.gjopen, B.cell_vol, B.diviterm, B.Dm, ei, c->c_mid,
                           ci, li, vi, co, lo, vo, c->fsa, B.gates, B.chan_mod, c->rates, c->fch,
                           c->flags, c->bad_step, c->step_ctr,
                           ob ? ob->fluxes_mem : nullptr, ob ? ob->fluxes_gj : nullptr,
                           ob ? ob->chan_flux : nullptr, ob ? ob->rate_NaKATP : nullptr, B.Phi_b},
               ExArgs{B.sq_ptr, B.sq_has_mem, B.sq_cluster, c->kcl, c->c_mid, c->fsa, c->fch, eo, c->u,
                      c->prm.flux_mode == BT_FLUX_FP32 ? 1 : 0, c->flags,
                      c->step_ctr, count_step, (1u << BT_MAX_STREAMS) - 1u, c->sq_live, c->n_live}};
    if (fused) {
        // the env kernel does the exchange: the surface charge ping-pongs
        // with the env state (the kernel reads u on a halo while it writes its
        // own squares' next value)
        a.field.u = parity ? c->u_b : c->u;
        a.field.u_out = parity ? c->u : c->u_b;
        a.field.sq_ptr = B.sq_ptr;
        a.field.has_mem = B.sq_has_mem;
        a.field.fsa = c->fsa;
        a.field.fsa32 = a.ex.fsa32;
        a.field.step_ctr = c->step_ctr;
        a.field.count_step = count_step;
        a.field.ctr_mask = a.ex.ctr_mask;
    }
    if (G > 1) {
        // the group's slices of the work lists; its own step counter (group 0
        // also advances the counters of groups this split does not use, so
        // that every counter holds the step count when a later split reads it)
        const Span sp = group_span(c, G, g);
        a.field.env_tiles += 3 * sp.e0;
        a.mem.tile_ptr += sp.t0;
        a.mem.tile_cluster += sp.t0;
        a.mem.step_ctr += g;
        a.ex.sq_live += sp.l0;
        a.ex.n_live = sp.nl;
        a.ex.ctr_mask = (g == 0) ? (((1u << BT_MAX_STREAMS) - 1u) & ~((1u << G) - 2u)) : (1u << g);
        a.field.ctr_mask = a.ex.ctr_mask;
    }
    return a;
}

// One step on the stream.
// which: -1 all, 0 membrane/cell, 1 env field+diffusion, 2 exchange, 3 field only
static void enqueue_step(bt_ctx* c, int parity, int which = -1, int G = 1, int g = 0) {
    const bt_obs* ob = c->capture;
    const bool fused = c->tune.exchange_form == BT_EXCHANGE_FUSED;
    const StepArgs a = step_args(c, parity, ob, which < 0 ? 1 : 0, G, g, fused);
    const Span sp = group_span(c, G, g);
    hipStream_t st = (G > 1) ? c->gstream[g] : c->stream;
    if (fused) {
        // membranes first (they read the pre-diffusion env only), then the
        // env kernel with the exchange
        if (which < 0 || which == 0) launch_mem_cell(c, a.mem, st, sp.nt);
        if (which < 0 || which == 1) launch_field_args(c, true, a.field, st, sp.ne);
    }
    if (!fused && (which < 0 || which == 1)) launch_field_args(c, true, a.field, st, sp.ne);
    if (!fused && (which < 0 || which == 0)) launch_mem_cell(c, a.mem, st, sp.nt);
    if (!fused && (which < 0 || which == 2)) {
        const int nl = sp.nl > 0 ? sp.nl : 1;    // block 0 also counts the step
        if (c->tune.exchange_form == 0)
            hipLaunchKernelGGL(k_env_exchange, dim3((nl + EXQ - 1) / EXQ), dim3(EXQ * c->prm.n_ions), 0, st, a.ex,
                               c->kp);
        else
            hipLaunchKernelGGL(k_env_exchange_sq, dim3(grid_blocks(nl)), dim3(BLOCK), 0, st, a.ex, c->kp);
    }
    if (which == 3) launch_field(c, false);
    if ((which < 0 || which == 4) && c->net && c->c_net)
        hipLaunchKernelGGL(k_network, dim3((c->prm.n_cells + NET_CELLS - 1) / NET_CELLS), dim3(NET_BLOCK),
                           net_lds_bytes(c->net_species, c->net_rates), c->stream, (const KNet*)c->net, c->c_net, c->prm.n_cells, c->prm.dt, c->flags,
                           c->bad_step, (const long long*)c->step_ctr, which < 0 ? 1 : 0);
}

// The persistent step kernel of this problem, or nullptr where it has no
// instantiation: five or six ions, no reaction network (the network keeps its
// own kernel).
static const void* persistent_kernel(bt_ctx* c) {
    if (c->net && c->c_net) return nullptr;
    const bool f32 = c->prm.flux_mode == BT_FLUX_FP32;
    const bool rates = c->prm.n_chan > 0;
    switch (c->prm.n_ions) {
        case 5:
            if (f32) return rates ? (const void*)&k_step_persistent<5, float, true> : (const void*)&k_step_persistent<5, float, false>;
            return rates ? (const void*)&k_step_persistent<5, double, true> : (const void*)&k_step_persistent<5, double, false>;
        case 6:
            if (f32) return rates ? (const void*)&k_step_persistent<6, float, true> : (const void*)&k_step_persistent<6, float, false>;
            return rates ? (const void*)&k_step_persistent<6, double, true> : (const void*)&k_step_persistent<6, double, false>;
        default:
            return nullptr;
    }
}

// blocks of a cooperative launch of the persistent kernel: as many as stay
// resident together (0 where the kernel or cooperative launches are unavailable)
static int persistent_grid(bt_ctx* c) {
    if (c->pers_grid >= 0) return c->pers_grid;
    c->pers_grid = 0;
    const void* k = persistent_kernel(c);
    int coop = 0, cus = 0, per_cu = 0;
    if (!k || hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device) != hipSuccess || !coop)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, PB, 0) != hipSuccess || per_cu < 1) return 0;
    c->pers_grid = per_cu * cus;
    return c->pers_grid;
}

// The blocks a persistent launch uses: as many as the largest phase has work
// items for (env (tile, ion group) items, cell tiles at PW per block, squares
// with membranes and (channel, cell) gating items at PB per block), rounded up
// to a multiple of the XCD count, and at most what stays resident. Every block
// arrives at each grid barrier, so a small problem waits on few arrivals.
static int persistent_launch_grid(bt_ctx* c) {
    const int cap = persistent_grid(c);
    if (cap < 1) return cap;
    const long long groups = (c->prm.n_ions + c->tune.env_ions_per_block - 1) / c->tune.env_ions_per_block;
    long long need = (long long)c->n_env_tiles * groups;
    need = std::max(need, (long long)(c->n_tiles + PW - 1) / PW);
    need = std::max(need, (long long)(c->n_live + PB - 1) / PB);
    need = std::max(need, ((long long)c->prm.n_chan * c->prm.n_cells + PB - 1) / PB);
    need = ((std::max(need, 1LL) + N_XCD - 1) / N_XCD) * N_XCD;
    return (int)std::min<long long>(need, cap);
}

// n_steps (even) steps in cooperative launches of the persistent kernel
static int launch_persistent(bt_ctx* c, int n_steps) {
    const void* k = persistent_kernel(c);
    const int grid = persistent_launch_grid(c);
    if (!k || grid < 1) return set_err(c, BT_EINVAL, "no persistent step kernel for this problem");
    PersistArgs pa;
    for (int par = 0; par < 2; ++par) {
        const StepArgs a = step_args(c, par, nullptr, 1);
        pa.field[par] = a.field;
        pa.mem[par] = a.mem;
        pa.ex[par] = a.ex;
    }
    pa.bar = c->bar;
    pa.n_env_tiles = c->n_env_tiles;
    pa.n_groups = (c->prm.n_ions + c->tune.env_ions_per_block - 1) / c->tune.env_ions_per_block;
    pa.n_tiles = c->n_tiles;
    for (int done = 0; done < n_steps;) {
        // at most BT_PERSIST_STEPS steps per launch
        pa.n_steps = std::min(n_steps - done, BT_PERSIST_STEPS);
        HIP_TRY(c, hipMemsetAsync(c->bar, 0, 2 * sizeof(unsigned), c->stream));
        void* args[] = {(void*)&pa, (void*)&c->kp};
        HIP_TRY(c, hipLaunchCooperativeKernel(k, dim3(grid), dim3(PB), args, 0, c->stream));
        done += pa.n_steps;
    }
    return BT_OK;
}

static void drop_graphs(bt_ctx* c) {
    for (int g = 0; g < BT_MAX_STREAMS; ++g) {
        if (c->gexec[g]) { (void)hipGraphExecDestroy(c->gexec[g]); c->gexec[g] = nullptr; }
        if (c->graph[g]) { (void)hipGraphDestroy(c->graph[g]); c->graph[g] = nullptr; }
        if (c->gexec_n[g]) { (void)hipGraphExecDestroy(c->gexec_n[g]); c->gexec_n[g] = nullptr; }
        if (c->graph_n[g]) { (void)hipGraphDestroy(c->graph_n[g]); c->graph_n[g] = nullptr; }
    }
}

// whether the problem can be stepped in G stream groups
static bool streams_allowed(const bt_ctx* c, int G) {
    return G >= 1 && G <= BT_MAX_STREAMS && (G == 1 || ((int)c->clusters.size() >= G && !c->cl_tile0.empty() &&
                                                        c->prm.n_chan == 0 && !(c->net && c->c_net)));
}

// the exchange fused into the env kernel: no membrane law may read the
// post-diffusion env (Ca pump, channels), and one block must hold every ion
// of a square (its charge)
static bool fused_allowed(const bt_ctx* c, const bt_tuning* t) {
    return c->prm.n_chan == 0 && c->prm.i_ca < 0 && t->env_ions_per_block == c->prm.n_ions;
}

// the stream groups of the current tuning (1: the whole problem on the caller's stream)
static int n_stream_groups(const bt_ctx* c) { return c->tune.persistent ? 1 : std::max(1, (int)c->tune.streams); }

// the groups' streams and the fork/join events, created on first use
static int ensure_group_streams(bt_ctx* c, int G) {
    c->gstream[0] = c->stream;
    if (G > 1 && !c->ev_fork) HIP_TRY(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    for (int g = 1; g < G; ++g) {
        if (!c->gstream[g]) HIP_TRY(c, hipStreamCreateWithFlags(&c->gstream[g], hipStreamNonBlocking));
        if (!c->ev_join[g]) HIP_TRY(c, hipEventCreateWithFlags(&c->ev_join[g], hipEventDisableTiming));
    }
    return BT_OK;
}

// Capture and instantiate the step graphs (one pair, and tune.graph_pairs
// pairs when that is more than one) of every stream group, each on its
// group's stream, if they do not exist yet.
static int capture_graphs(bt_ctx* c) {
    const int G = n_stream_groups(c);
    int rc = ensure_group_streams(c, G);
    if (rc != BT_OK) return rc;
    auto capture = [&](int g, int npairs, hipGraph_t* gr, hipGraphExec_t* x) -> int {
        hipStream_t st = (G > 1) ? c->gstream[g] : c->stream;
        HIP_TRY(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < npairs; ++q) {
            enqueue_step(c, 0, -1, G, g);
            enqueue_step(c, 1, -1, G, g);
        }
        hipError_t e = hipStreamEndCapture(st, gr);
        if (e != hipSuccess) return set_err(c, BT_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
        HIP_TRY(c, hipGraphInstantiate(x, *gr, nullptr, nullptr, 0));
        HIP_TRY(c, hipGraphUpload(*x, st));
        return BT_OK;
    };
    for (int g = 0; g < G; ++g) {
        if (!c->gexec[g]) {
            rc = capture(g, 1, &c->graph[g], &c->gexec[g]);
            if (rc != BT_OK) return rc;
        }
        if (c->tune.graph_pairs > 1 && !c->gexec_n[g]) {
            rc = capture(g, c->tune.graph_pairs, &c->graph_n[g], &c->gexec_n[g]);
            if (rc != BT_OK) return rc;
        }
    }
    return BT_OK;
}

// Replay the step graphs for `pairs` pairs of steps: tune.graph_pairs pairs
// per launch of the long graph, single pairs for the rest. With stream groups
// the other groups' streams first wait for the caller's stream, and the
// caller's stream then waits for all of them.
static int launch_graphs(bt_ctx* c, int pairs) {
    const int G = n_stream_groups(c);
    if (G > 1) {
        HIP_TRY(c, hipEventRecord(c->ev_fork, c->stream));
        for (int g = 1; g < G; ++g) HIP_TRY(c, hipStreamWaitEvent(c->gstream[g], c->ev_fork, 0));
    }
    const int gp = c->tune.graph_pairs;
    for (int g = 0; g < G; ++g) {
        hipStream_t st = (G > 1) ? c->gstream[g] : c->stream;
        int done = 0;
        if (c->gexec_n[g])
            for (; done + gp <= pairs; done += gp) HIP_TRY(c, hipGraphLaunch(c->gexec_n[g], st));
        for (; done < pairs; ++done) HIP_TRY(c, hipGraphLaunch(c->gexec[g], st));
    }
    if (G > 1)
        for (int g = 1; g < G; ++g) {
            HIP_TRY(c, hipEventRecord(c->ev_join[g], c->gstream[g]));
            HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ev_join[g], 0));
        }
    return BT_OK;
}

static int copy_back(bt_ctx* c) {
    const size_t nc = (size_t)c->prm.n_cells;
    const size_t ni = (size_t)c->prm.n_ions;
    HIP_TRY(c, hipMemcpyAsync(c->buf.c_cell, c->c_cell_b, ni * nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->buf.c_lag, c->c_lag_b, ni * nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->buf.vm_cell, c->vm_b, nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->buf.c_env, c->c_env_b, ni * (size_t)c->kp.n_sq * 8, hipMemcpyDeviceToDevice,
                              c->stream));
    if (c->tune.exchange_form == BT_EXCHANGE_FUSED)
        HIP_TRY(c, hipMemcpyAsync(c->u, c->u_b, (size_t)c->kp.n_sq * 8, hipMemcpyDeviceToDevice, c->stream));
    return BT_OK;
}

// scalar arguments of bt_create / bt_validate
static int check_params(const bt_params* p, const bt_cluster* cl, int32_t n_clusters, const bt_buffers* b) {
    if (p->n_ions < 1 || p->n_ions > BT_MAX_IONS) return set_err(nullptr, BT_EINVAL, "n_ions out of range");
    if (p->n_cells < 1 || p->n_mems < 1) return set_err(nullptr, BT_EINVAL, "empty cluster");
    if (n_clusters < 1) return set_err(nullptr, BT_EINVAL, "n_clusters < 1");
    {
        long long off = 0;
        for (int q = 0; q < n_clusters; ++q) {
            if (cl[q].grid_ny < 3 || cl[q].grid_nx < 3) return set_err(nullptr, BT_EINVAL, "grid must be at least 3x3");
            if (cl[q].sq_off != off) return set_err(nullptr, BT_EINVAL, "cluster grids must be consecutive");
            if (!(cl[q].delta > 0) || !(cl[q].mean_sa > 0) || !(cl[q].ko_env > 0))
                return set_err(nullptr, BT_EINVAL, "cluster constants must be positive");
            off += (long long)cl[q].grid_ny * cl[q].grid_nx;
        }
        if (off != p->n_squares || off > (1LL << 30)) return set_err(nullptr, BT_EINVAL, "n_squares mismatch");
    }
    if (p->i_na >= p->n_ions || p->i_k >= p->n_ions || p->i_ca >= p->n_ions)
        return set_err(nullptr, BT_EINVAL, "ion slot out of range");
    if ((p->i_na < 0) != (p->i_k < 0)) return set_err(nullptr, BT_EINVAL, "NaK pump needs both Na and K");
    if (p->flux_mode != BT_FLUX_FP64 && p->flux_mode != BT_FLUX_FP32)
        return set_err(nullptr, BT_EINVAL, "unknown flux_mode");
    if (b->n_tiles < 1) return set_err(nullptr, BT_EINVAL, "n_tiles < 1");
    if (b->tile_m != BT_TILE_MEMBRANES) return set_err(nullptr, BT_EINVAL, "tiles must hold <= 64 membranes");
    if (p->n_chan < 0 || p->n_chan > BT_MAX_CHAN) return set_err(nullptr, BT_EINVAL, "n_chan out of range");
    for (int c = 0; c < p->n_chan; ++c) {
        if (p->chan_kind[c] != BT_CHAN_NAV1P2 && p->chan_kind[c] != BT_CHAN_KV1P1 &&
            p->chan_kind[c] != BT_CHAN_CAV1P2)
            return set_err(nullptr, BT_EINVAL, "unknown channel kind");
        if (p->chan_ion[c] < 0 || p->chan_ion[c] >= p->n_ions)
            return set_err(nullptr, BT_EINVAL, "channel ion slot out of range");
        for (int d = 0; d < c; ++d)
            if (p->chan_ion[d] == p->chan_ion[c]) return set_err(nullptr, BT_EINVAL, "two channels on one ion");
    }
    if (p->n_chan > 0 && (!b->gates || !b->chan_mod))
        return set_err(nullptr, BT_EINVAL, "channel buffers missing");
    return BT_OK;
}

// the index arrays of a batch, from host copies (bt_validate's contract)
static int check_geometry(const bt_params* p, const bt_cluster* cl, int32_t n_clusters, const bt_buffers* h) {
    const int C = p->n_cells, M = p->n_mems, S = p->n_squares, T = h->n_tiles;
    const int32_t* tp = h->tile_ptr; const int32_t* mp = h->mem_ptr; const int32_t* tcl = h->tile_cluster;
    auto bad = [&](const std::string& msg) { return set_err(nullptr, BT_EINVAL, msg); };
    // tiles and the cell CSR (k_mem_cell indexes LDS by a tile's membranes and cells)
    if (tp[0] != 0 || tp[T] != C || mp[0] != 0 || mp[C] != M)
        return bad("tile_ptr/mem_ptr do not cover the cells and membranes");
    for (int q = 0; q < C; ++q)
        if (mp[q + 1] <= mp[q]) return bad("a cell has no membranes (mem_ptr must increase)");
    std::vector<int32_t> cell_cl((size_t)C);
    for (int t = 0; t < T; ++t) {
        const int a0 = tp[t], a1 = tp[t + 1];
        if (a1 <= a0 || a1 - a0 > BT_TILE_CELLS || a1 > C || mp[a1] - mp[a0] > BT_TILE_MEMBRANES)
            return bad("tile " + std::to_string(t) + " exceeds the tile limits");
        if (tcl[t] < 0 || tcl[t] >= n_clusters) return bad("tile_cluster out of range");
        for (int q = a0; q < a1; ++q) cell_cl[q] = tcl[t];
    }
    // squares: cluster of each square from the cluster grids
    for (int q = 0; q < n_clusters; ++q) {
        const long long k0 = cl[q].sq_off, k1 = k0 + (long long)cl[q].grid_ny * cl[q].grid_nx;
        for (long long k = k0; k < k1; ++k)
            if (h->sq_cluster[k] != q) return bad("sq_cluster does not match the cluster grids");
    }
    // membranes: owner cell, then gap-junction partner and square within the membrane's cluster
    for (int c = 0; c < C; ++c)
        for (int m = mp[c]; m < mp[c + 1]; ++m)
            if (h->mem_cell[m] != c) return bad("mem_cell is not the CSR owner of membrane " + std::to_string(m));
    for (int c = 0; c < C; ++c)
        for (int m = mp[c]; m < mp[c + 1]; ++m) {
            const int nm = h->nn_mem[m];
            if (nm < 0 || nm >= M) return bad("nn_mem out of range at membrane " + std::to_string(m));
            if (h->nn_cell[m] != h->mem_cell[nm]) return bad("nn_cell is not the cell of nn_mem at membrane " + std::to_string(m));
            if (cell_cl[h->nn_cell[m]] != cell_cl[c]) return bad("a gap-junction partner lies in another cluster");
            const int k = h->mem_sq[m];
            if (k < 0 || k >= S) return bad("mem_sq out of range at membrane " + std::to_string(m));
            if (h->sq_cluster[k] != cell_cl[c]) return bad("mem_sq maps a membrane to another cluster's grid");
        }
    // grid-side CSR: every membrane once, under its own square; mem_pos its inverse
    const int32_t* sp = h->sq_ptr;
    if (sp[0] != 0 || sp[S] != M) return bad("sq_ptr does not cover the membranes");
    std::vector<uint8_t> seen((size_t)M, 0);
    for (int k = 0; k < S; ++k) {
        if (sp[k + 1] < sp[k]) return bad("sq_ptr decreases");
        for (int q = sp[k]; q < sp[k + 1]; ++q) {
            const int m = h->sq_mems[q];
            if (m < 0 || m >= M || seen[m]) return bad("sq_mems is not a permutation of the membranes");
            seen[m] = 1;
            if (h->mem_sq[m] != k) return bad("sq_mems lists a membrane under another square");
            if (h->mem_pos[m] != q) return bad("mem_pos is not the inverse of sq_mems");
        }
    }
    return BT_OK;
}

// the FP32-flux path's elementary functions over n floats (test hook)
__global__ void k_fp32_math(const float* x, const float* y, float* out, int n, int fn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (fn == 0) out[t] = exp32(x[t]);
    else if (fn == 1) out[t] = expm1_32(x[t]);
    else out[t] = dv(x[t], y[t]);
}

__global__ void k_fp64_math(const double* x, const double* y, double* out, int n, int fn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (fn == 0) out[t] = exp64(x[t]);
    else if (fn == 1) out[t] = expm1_64(x[t]);
    else out[t] = dv(x[t], y[t]);
}

extern "C" {

int bt_version(void) { return 301; }

const char* bt_last_error(const bt_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int64_t bt_steps_done(const bt_ctx* ctx) { return ctx ? ctx->steps_done : -1; }

// The per-cluster constants read from the caller's Dm, gjopen and D_env: a
// value that is the same over all of a cluster's membranes (squares) becomes
// a constant of the cluster table, so the kernels skip its loads; a still ion
// (no membrane flux anywhere in the cluster: Dm zero, no pump or channel
// carrying it) is marked. Into kcl_host; the caller uploads the table.
static int derive_uniform(bt_ctx* c) {
    const bt_params* p = &c->prm;
    const int nq = (int)c->clusters.size();
    std::vector<double> dm((size_t)p->n_ions * p->n_mems), gj((size_t)p->n_mems), de((size_t)p->n_ions * p->n_squares);
    std::vector<double> pb(c->buf.Phi_b ? (size_t)p->n_squares : 0);
    hipError_t e = hipMemcpy(dm.data(), c->buf.Dm, dm.size() * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(gj.data(), c->buf.gjopen, gj.size() * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(de.data(), c->buf.D_env, de.size() * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && !pb.empty()) e = hipMemcpy(pb.data(), c->buf.Phi_b, pb.size() * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return set_err(c, BT_EHIP, std::string("reading Dm, gjopen, D_env, Phi_b: ") + hipGetErrorString(e));
    c->any_phib = false;
    for (int q = 0; q < nq; ++q) {
        KCluster& K = c->kcl_host[q];
        const int lo = c->cl_mem_lo[q], hi = c->cl_mem_hi[q];
        for (int i = 0; i < BT_MAX_IONS; ++i) K.dm_u[i] = K.denv_u[i] = NAN;
        K.gjo_u = NAN;
        K.still_mem = 0;
        if (hi > lo) {
            for (int i = 0; i < p->n_ions; ++i) {
                const double* d = dm.data() + (size_t)i * p->n_mems;
                bool same = true;
                for (int m = lo; m < hi && same; ++m) same = (d[m] == d[lo]);
                K.dm_u[i] = same ? d[lo] : NAN;
            }
            bool same = true;
            for (int m = lo; m < hi && same; ++m) same = (gj[m] == gj[lo]);
            K.gjo_u = same ? gj[lo] : NAN;
            for (int i = 0; i < p->n_ions; ++i)
                if (K.dm_u[i] == 0.0 && i != c->kp.i_na && i != c->kp.i_k && i != c->kp.i_ca && c->kp.ion_chan[i] < 0)
                    K.still_mem |= 1u << i;
        }
        const size_t k0 = (size_t)c->clusters[q].sq_off;
        const size_t k1 = k0 + (size_t)c->clusters[q].grid_ny * c->clusters[q].grid_nx;
        for (int i = 0; i < p->n_ions; ++i) {
            const double* d = de.data() + (size_t)i * p->n_squares;
            bool same = true;
            for (size_t k = k0; k < k1 && same; ++k) same = (d[k] == d[k0]);
            K.denv_u[i] = same ? d[k0] : NAN;
        }
        K.phib = 0;
        for (size_t k = k0; k < k1 && !pb.empty() && !K.phib; ++k) K.phib = (pb[k] != 0.0);
        if (K.phib) K.still_mem |= PHIB_BIT;
        if (K.phib && p->n_chan > 0)
            return set_err(c, BT_EINVAL, "an applied boundary voltage (Phi_b) with voltage-gated channels is not "
                                         "supported (their gating rates are per cell)");
        c->any_phib = c->any_phib || K.phib;
    }
    return BT_OK;
}

int bt_refresh_params(bt_ctx* c) {
    if (!c) return BT_EINVAL;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int g = 1; g < BT_MAX_STREAMS; ++g)
        if (c->gstream[g]) HIP_TRY(c, hipStreamSynchronize(c->gstream[g]));
    int rc = derive_uniform(c);
    if (rc != BT_OK) return rc;
    HIP_TRY(c, hipMemcpy(c->kcl, c->kcl_host.data(), sizeof(KCluster) * c->kcl_host.size(), hipMemcpyHostToDevice));
    return BT_OK;
}

int bt_validate(const bt_params* p, const bt_cluster* cl, int32_t n_clusters, const bt_buffers* h) {
    if (!p || !cl || !h) return set_err(nullptr, BT_EINVAL, "null argument");
    int rc = check_params(p, cl, n_clusters, h);
    if (rc != BT_OK) return rc;
    const void* req[] = {h->mem_ptr, h->mem_cell, h->nn_mem, h->nn_cell, h->mem_sq, h->sq_ptr, h->sq_mems,
                         h->mem_pos, h->tile_ptr, h->tile_cluster, h->sq_cluster};
    for (const void* q : req)
        if (!q) return set_err(nullptr, BT_EINVAL, "null geometry array");
    return check_geometry(p, cl, n_clusters, h);
}

int bt_create(bt_ctx** out, const bt_params* p, const bt_cluster* cl, int32_t n_clusters,
              const bt_buffers* b, void* stream) {
    if (!out || !p || !b || !cl) return set_err(nullptr, BT_EINVAL, "null argument");
    *out = nullptr;
    int rc0 = check_params(p, cl, n_clusters, b);
    if (rc0 != BT_OK) return rc0;
    const void* req[] = {b->mem_ptr, b->mem_cell, b->nn_mem, b->nn_cell, b->mem_sq, b->mem_pos, b->sq_ptr,
                         b->tile_ptr, b->tile_cluster, b->sq_cluster, b->mem_sa, b->gjopen, b->cell_vol, b->diviterm, b->Dm, b->D_env,
                         b->sq_has_mem, b->c_cell, b->c_lag, b->vm_cell, b->c_env, b->E_x, b->E_y, b->v_env};
    for (const void* q : req)
        if (!q) return set_err(nullptr, BT_EINVAL, "null device buffer");
    // every index array the kernels gather or scatter through, checked on host copies
    const size_t NM = (size_t)p->n_mems, NC = (size_t)p->n_cells, NS = (size_t)p->n_squares;
    std::vector<int32_t> tp((size_t)b->n_tiles + 1), mp(NC + 1), tcl((size_t)b->n_tiles), sp(NS + 1),
        mcell(NM), nnm(NM), nnc(NM), msq(NM), sqm(NM), mpos(NM), sqc(NS);
    std::vector<uint8_t> hm(NS);
    {
        struct { void* dst; const void* src; size_t bytes; } cp[] = {
            {tp.data(), b->tile_ptr, tp.size() * 4}, {mp.data(), b->mem_ptr, mp.size() * 4},
            {tcl.data(), b->tile_cluster, tcl.size() * 4}, {sp.data(), b->sq_ptr, sp.size() * 4},
            {mcell.data(), b->mem_cell, NM * 4}, {nnm.data(), b->nn_mem, NM * 4}, {nnc.data(), b->nn_cell, NM * 4},
            {msq.data(), b->mem_sq, NM * 4}, {sqm.data(), b->sq_mems, NM * 4}, {mpos.data(), b->mem_pos, NM * 4},
            {sqc.data(), b->sq_cluster, NS * 4}, {hm.data(), b->sq_has_mem, NS}};
        for (auto& q : cp) {
            hipError_t e = hipMemcpy(q.dst, q.src, q.bytes, hipMemcpyDeviceToHost);
            if (e != hipSuccess)
                return set_err(nullptr, BT_EHIP, std::string("reading the geometry: ") + hipGetErrorString(e));
        }
        bt_buffers h = *b;
        h.tile_ptr = tp.data(); h.mem_ptr = mp.data(); h.tile_cluster = tcl.data(); h.sq_ptr = sp.data();
        h.mem_cell = mcell.data(); h.nn_mem = nnm.data(); h.nn_cell = nnc.data(); h.mem_sq = msq.data();
        h.sq_mems = sqm.data(); h.mem_pos = mpos.data(); h.sq_cluster = sqc.data();
        int rc = check_geometry(p, cl, n_clusters, &h);
        if (rc != BT_OK) return rc;
    }
    // the squares k_env_exchange updates: those with membranes or a surface
    // charge (the others' exchange is the identity and their u stays 0)
    std::vector<int32_t> sq_live;
    std::vector<uint8_t> live_flag(NS, 0);
    for (size_t k = 0; k < NS; ++k) {
        if (sp[k + 1] > sp[k] || hm[k]) {
            sq_live.push_back((int32_t)k);
            live_flag[k] = 1;
        }
    }

    bt_ctx* c = new bt_ctx();
    c->prm = *p;
    c->buf = *b;
    c->n_tiles = b->n_tiles;
    c->stream = (hipStream_t)stream;
    derive_params(*p, c->kp);
    c->clusters.assign(cl, cl + n_clusters);
    std::vector<int> tiles;
    for (int q = 0; q < n_clusters; ++q) {
        c->kcl_host.push_back(derive_cluster(*p, cl[q]));
        for (int r0 = 0; r0 < cl[q].grid_ny; r0 += TY)
            for (int c0 = 0; c0 < cl[q].grid_nx; c0 += TX) {
                tiles.push_back(q);
                tiles.push_back(r0);
                tiles.push_back(c0);
            }
    }
    c->n_env_tiles = (int)(tiles.size() / 3);
    {
        // per cluster, the start of its cell tiles, env tiles, live squares and
        // cells, for stream groups; left empty unless each list is in cluster order
        const int nq = n_clusters, nt = b->n_tiles, ne = c->n_env_tiles, nl = (int)sq_live.size();
        bool ordered = true;
        for (int t = 1; t < nt && ordered; ++t) ordered = tcl[t] >= tcl[t - 1];
        for (int l = 1; l < nl && ordered; ++l) ordered = sqc[sq_live[l]] >= sqc[sq_live[l - 1]];
        if (ordered) {
            auto starts = [&](auto key, int n) {
                std::vector<int> out(nq + 1);
                int i = 0;
                for (int q = 0; q <= nq; ++q) {
                    while (i < n && key(i) < q) ++i;
                    out[q] = i;
                }
                return out;
            };
            c->cl_tile0 = starts([&](int t) { return tcl[t]; }, nt);
            c->cl_env0 = starts([&](int t) { return tiles[3 * t]; }, ne);
            c->cl_live0 = starts([&](int l) { return sqc[sq_live[l]]; }, nl);
            c->cl_cell0.assign(nq + 1, 0);
            for (int q = 0; q < nq; ++q)
                c->cl_cell0[q + 1] = c->cl_cell0[q] + (tp[c->cl_tile0[q + 1]] - tp[c->cl_tile0[q]]);
        }
    }
    {
        // each cluster's membrane range, for the per-cluster constants
        c->cl_mem_lo.assign(n_clusters, p->n_mems);
        c->cl_mem_hi.assign(n_clusters, 0);
        for (int t = 0; t < b->n_tiles; ++t) {
            c->cl_mem_lo[tcl[t]] = std::min(c->cl_mem_lo[tcl[t]], mp[tp[t]]);
            c->cl_mem_hi[tcl[t]] = std::max(c->cl_mem_hi[tcl[t]], mp[tp[t + 1]]);
        }
        const int rc = derive_uniform(c);
        if (rc != BT_OK) {
            const std::string msg = c->err;
            delete c;
            return set_err(nullptr, rc, msg);
        }
    }
    (void)hipGetDevice(&c->device);
    const size_t nc = p->n_cells, nm = p->n_mems, ni = p->n_ions, ns = (size_t)p->n_squares;
    hipError_t e = hipSuccess;
    auto alloc = [&](void** ptr, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(ptr, bytes < 256 ? 256 : bytes);
    };
    alloc((void**)&c->c_cell_b, ni * nc * 8);
    alloc((void**)&c->c_lag_b, ni * nc * 8);
    alloc((void**)&c->vm_b, nc * 8);
    alloc((void**)&c->c_mid, ni * ns * 8);
    alloc((void**)&c->fsa, ni * nm * 8);
    alloc((void**)&c->fch, (p->n_chan > 0 ? (size_t)p->n_chan : 1) * nm * 8);
    alloc((void**)&c->rates, (p->n_chan > 0 ? (size_t)p->n_chan : 1) * 4 * nc * 8);
    alloc((void**)&c->u, ns * 8);
    alloc((void**)&c->u_b, ns * 8);
    alloc((void**)&c->flags, 256);
    alloc((void**)&c->bad_step, 256);
    alloc((void**)&c->step_ctr, 256);
    alloc((void**)&c->bar, 256);
    alloc((void**)&c->kcl, sizeof(KCluster) * (size_t)n_clusters);
    alloc((void**)&c->env_tiles, sizeof(int) * tiles.size());
    alloc((void**)&c->c_env_b, ni * ns * 8);
    alloc((void**)&c->sq_live, sizeof(int32_t) * (sq_live.empty() ? 1 : sq_live.size()));
    alloc((void**)&c->live_flag, ns);
    c->n_live = (int)sq_live.size();
    c->tune.mem_waves = (c->n_tiles < BT_MEM_WG_TILES) ? 2 : 1;
    c->tune.env_ions_per_block = (c->n_env_tiles < ENV_SPLIT_TILES) ? std::min(2, p->n_ions) : p->n_ions;
    c->tune.exchange_form = (c->n_live < EX_SPLIT_SQUARES) ? 0 : 1;
    c->tune.graph_pairs = BT_GRAPH_PAIRS;
    c->tune.persistent = 0;
    c->tune.streams = 1;
    // large problems whose membrane laws read the pre-diffusion env only:
    // the exchange in the env kernel (measured faster on the 500k-cell batch)
    if (c->tune.exchange_form == 1 && fused_allowed(c, &c->tune)) c->tune.exchange_form = BT_EXCHANGE_FUSED;
    if (e == hipSuccess && !sq_live.empty())
        e = hipMemcpy(c->sq_live, sq_live.data(), sizeof(int32_t) * sq_live.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->live_flag, live_flag.data(), ns, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->kcl, c->kcl_host.data(), sizeof(KCluster) * (size_t)n_clusters, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->env_tiles, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(c->flags, 0, 256, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->bad_step, 0xff, 256, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->step_ctr, 0, 256, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->bar, 0, 256, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->u, 0, ns * 8, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->u_b, 0, ns * 8, c->stream);    // u of squares without membranes
    if (e == hipSuccess) e = hipMemsetAsync(c->fsa, 0, ni * nm * 8, c->stream);   // still rows stay zero
    if (e == hipSuccess)
        e = hipMemsetAsync(c->rates, 0, (p->n_chan > 0 ? (size_t)p->n_chan : 1) * 4 * nc * 8, c->stream);
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
    drop_graphs(c);
    for (int g = 1; g < BT_MAX_STREAMS; ++g) {
        if (c->gstream[g]) (void)hipStreamDestroy(c->gstream[g]);
        if (c->ev_join[g]) (void)hipEventDestroy(c->ev_join[g]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    void* ptrs[] = {c->c_cell_b, c->c_lag_b, c->vm_b, c->c_mid, c->fsa, c->fch, c->u, c->u_b, c->rates, c->c_env_b,
                    c->sq_live, c->live_flag, c->flags, c->bad_step, c->step_ctr, c->kcl, c->env_tiles,
                    c->vm_prev, c->net, c->bar, c->phib_stage, c->phib_prev};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    for (void* q : c->obs_alloc) (void)hipFree(q);
    delete c;
    return BT_OK;
}

int bt_update_v(bt_ctx* c) {
    if (!c) return BT_EINVAL;
    const int nsq = c->kp.n_sq;
    hipLaunchKernelGGL(k_cell_vm, dim3(grid_blocks(c->kp.n_cells)), dim3(BLOCK), 0, c->stream,
                       (const double*)c->buf.c_cell, c->buf.diviterm, c->buf.vm_cell, c->kp);
    hipLaunchKernelGGL(k_env_charge, dim3(grid_blocks(nsq)), dim3(BLOCK), 0, c->stream,
                       (const double*)c->buf.c_env, c->buf.sq_has_mem, c->buf.sq_cluster,
                       (const KCluster*)c->kcl, c->u, c->kp);
    launch_field(c, false);
    HIP_TRY(c, hipGetLastError());
    return BT_OK;
}

// A staged Phi_b becomes the one in effect (update_V's Phi_b at the end of the
// step the change fires in), and the per-cluster flags follow it.
static int apply_staged_boundary(bt_ctx* c) {
    HIP_TRY(c, hipMemcpyAsync(c->buf.Phi_b, c->phib_stage, (size_t)c->prm.n_squares * 8, hipMemcpyDeviceToDevice,
                              c->stream));
    c->phib_staged = false;
    return bt_refresh_params(c);
}

int bt_stage_boundary(bt_ctx* c, const double* Phi_b_next) {
    if (!c || !Phi_b_next) return BT_EINVAL;
    if (!c->buf.Phi_b) return set_err(c, BT_EINVAL, "the context was created without a Phi_b buffer");
    const size_t n = (size_t)c->prm.n_squares;
    for (size_t k = 0; k < n; ++k) {
        if (!std::isfinite(Phi_b_next[k])) return set_err(c, BT_EINVAL, "Phi_b is not finite");
        if (c->prm.n_chan > 0 && Phi_b_next[k] != 0.0)
            return set_err(c, BT_EINVAL, "an applied boundary voltage (Phi_b) with voltage-gated channels is not "
                                         "supported (their gating rates are per cell)");
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (!c->phib_stage) HIP_TRY(c, hipMalloc((void**)&c->phib_stage, n * 8));
    HIP_TRY(c, hipMemcpy(c->phib_stage, Phi_b_next, n * 8, hipMemcpyHostToDevice));
    c->phib_staged = true;
    return BT_OK;
}

int bt_step(bt_ctx* c, int32_t n_steps, int32_t use_graph) {
    if (!c) return BT_EINVAL;
    if (n_steps < 0) return set_err(c, BT_EINVAL, "n_steps < 0");
    if (n_steps == 0) return BT_OK;
    if (c->phib_staged) {
        // the step the change fires in runs with the Phi_b in effect; the
        // staged one holds from its end on
        enqueue_step(c, 0);
        int rc = copy_back(c);
        if (rc == BT_OK) rc = apply_staged_boundary(c);
        if (rc != BT_OK) return rc;
        c->steps_done += 1;
        if (n_steps == 1) {
            enqueue_step(c, 0, 3);
            HIP_TRY(c, hipGetLastError());
            return BT_OK;
        }
        return bt_step(c, n_steps - 1, use_graph);
    }
    int pairs = n_steps / 2;
    if (c->tune.persistent && pairs > 0) {
        // every pair of steps in cooperative launches of the persistent kernel
        int rc = launch_persistent(c, 2 * pairs);
        if (rc != BT_OK) return rc;
    } else if (use_graph && pairs > 0) {
        // a graph of tune.graph_pairs pairs of steps for the bulk, one of a
        // single pair for the rest; both are captured and instantiated on the
        // first call that replays graphs, so later calls only launch them
        int rc = capture_graphs(c);
        if (rc != BT_OK) return rc;
        rc = launch_graphs(c, pairs);
        if (rc != BT_OK) return rc;
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
    // v_env and E_env of the final update_V, for the host
    enqueue_step(c, 0, 3);
    HIP_TRY(c, hipGetLastError());
    c->steps_done += n_steps;
    return BT_OK;
}

int bt_get_tuning(const bt_ctx* c, bt_tuning* out) {
    if (!c || !out) return BT_EINVAL;
    *out = c->tune;
    return BT_OK;
}

int bt_set_tuning(bt_ctx* c, const bt_tuning* t) {
    if (!c || !t) return BT_EINVAL;
    if (t->mem_waves != 1 && t->mem_waves != 2) return set_err(c, BT_EINVAL, "mem_waves must be 1 or 2");
    if (t->env_ions_per_block < 1 || t->env_ions_per_block > c->prm.n_ions)
        return set_err(c, BT_EINVAL, "env_ions_per_block must be in 1..n_ions");
    if (t->exchange_form < 0 || t->exchange_form > BT_EXCHANGE_FUSED)
        return set_err(c, BT_EINVAL, "exchange_form must be 0, 1 or 2");
    if (t->exchange_form == BT_EXCHANGE_FUSED && !fused_allowed(c, t))
        return set_err(c, BT_EINVAL, "the exchange fused into the env kernel needs no channels, no Ca pump and "
                                     "env_ions_per_block = n_ions");
    if (t->graph_pairs < 1 || t->graph_pairs > BT_MAX_GRAPH_PAIRS)
        return set_err(c, BT_EINVAL, "graph_pairs must be in 1..64");
    if (t->persistent != 0 && t->persistent != 1) return set_err(c, BT_EINVAL, "persistent must be 0 or 1");
    if (t->persistent && persistent_grid(c) < 1)
        return set_err(c, BT_EINVAL, "no persistent step kernel for this problem (it needs 5 or 6 ions, no "
                                     "reaction network, and cooperative launches)");
    if (t->streams < 1 || t->streams > BT_MAX_STREAMS) return set_err(c, BT_EINVAL, "streams must be in 1..4");
    if (t->streams > 1 && !streams_allowed(c, t->streams))
        return set_err(c, BT_EINVAL, "stream groups need at least that many clusters, work lists in cluster order, "
                                     "no channels and no reaction network");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int g = 1; g < BT_MAX_STREAMS; ++g)
        if (c->gstream[g]) HIP_TRY(c, hipStreamSynchronize(c->gstream[g]));
    c->tune = *t;
    drop_graphs(c);
    return BT_OK;
}

// average ms of one launch of a step kernel (bt_time_kernel's timing)
static int time_launches(bt_ctx* c, int which, int n, double* ms_out) {
    hipEvent_t e0, e1;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    enqueue_step(c, 0, which);   // warm-up launch
    HIP_TRY(c, hipEventRecord(e0, c->stream));
    for (int s = 0; s < n; ++s) enqueue_step(c, 0, which);
    HIP_TRY(c, hipEventRecord(e1, c->stream));
    HIP_TRY(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(c, hipGetLastError());
    *ms_out = (double)ms / n;
    return BT_OK;
}

// The state a step modifies in place, as (buffer, bytes) regions: the
// caller's cell and env state, fields and gates, the surface charge u, the
// substances, and the status words.
static std::vector<std::pair<void*, size_t>> state_regions(bt_ctx* c) {
    const size_t C = (size_t)c->prm.n_cells, M = (size_t)c->prm.n_mems, NS = (size_t)c->kp.n_sq;
    const size_t ni = (size_t)c->prm.n_ions;
    std::vector<std::pair<void*, size_t>> r = {
        {c->buf.c_cell, ni * C * 8}, {c->buf.c_lag, ni * C * 8}, {c->buf.vm_cell, C * 8},
        {c->buf.c_env, ni * NS * 8}, {c->buf.E_x, NS * 8}, {c->buf.E_y, NS * 8}, {c->buf.v_env, NS * 8},
        {c->u, NS * 8}, {c->flags, 256}, {c->bad_step, 256}, {c->step_ctr, 256}};
    if (c->prm.n_chan > 0) r.push_back({c->buf.gates, (size_t)c->prm.n_chan * 2 * M * 8});
    if (c->net && c->c_net) r.push_back({c->c_net, (size_t)c->net_species * C * 8});
    return r;
}

// copy the regions into save (back = false) or back out of it, synchronously
static hipError_t state_copy(bt_ctx* c, const std::vector<std::pair<void*, size_t>>& regions, char* save,
                             bool back) {
    size_t o = 0;
    for (const auto& r : regions) {
        hipError_t e = back ? hipMemcpyAsync(r.first, save + o, r.second, hipMemcpyDeviceToDevice, c->stream)
                            : hipMemcpyAsync(save + o, r.first, r.second, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) return e;
        o += (r.second + 255) / 256 * 256;
    }
    return hipStreamSynchronize(c->stream);
}

int bt_capture_graphs(bt_ctx* c) {
    if (!c) return BT_EINVAL;
    bool fresh = false;
    for (int g = 0; g < n_stream_groups(c); ++g)
        fresh = fresh || !c->gexec[g] || (c->tune.graph_pairs > 1 && !c->gexec_n[g]);
    int rc = capture_graphs(c);
    if (rc != BT_OK || !fresh) return rc;
    // replay each graph once from a saved copy of the state and put the state
    // back: the first launch of an instantiated graph carries one-time costs
    // (measured ~1.4 ms for eight pairs of steps on the 500k-cell batch) that
    // would otherwise fall into the caller's first steps
    const auto regions = state_regions(c);
    size_t sbytes = 0;
    for (const auto& r : regions) sbytes += (r.second + 255) / 256 * 256;
    char* save = nullptr;
    HIP_TRY(c, hipMalloc((void**)&save, sbytes));
    hipError_t e = state_copy(c, regions, save, false);
    if (e == hipSuccess) {
        rc = launch_graphs(c, c->tune.graph_pairs > 1 ? c->tune.graph_pairs + 1 : 1);
        if (rc == BT_OK) e = state_copy(c, regions, save, true);
    }
    (void)hipFree(save);
    if (rc != BT_OK) return rc;
    if (e != hipSuccess) return set_err(c, BT_EHIP, std::string("graph warm-up: ") + hipGetErrorString(e));
    return BT_OK;
}

int bt_autotune(bt_ctx* c, int32_t n_launches, bt_tuning* chosen, double* trial_ms, int32_t n_trial_ms) {
    if (!c || n_launches < 1 || (trial_ms && n_trial_ms < 0)) return BT_EINVAL;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    unsigned flags = 0;
    HIP_TRY(c, hipMemcpy(&flags, c->flags, sizeof(flags), hipMemcpyDeviceToHost));
    if (flags) return set_err(c, BT_EINVAL, "autotune needs a stable state (a NaN was flagged)");
    if (c->phib_staged)
        return set_err(c, BT_EINVAL, "autotune with a staged boundary change: tune before staging it");
    // every timed launch starts from the state as it is now: keep a copy of all
    // that a step modifies in place and put it back after each trial
    const auto regions = state_regions(c);
    size_t sbytes = 0;
    for (const auto& r : regions) sbytes += (r.second + 255) / 256 * 256;
    char* save = nullptr;
    HIP_TRY(c, hipMalloc((void**)&save, sbytes));
    hipError_t e = state_copy(c, regions, save, false);
    const bt_tuning prev = c->tune;
    // the per-kernel trials run the three-kernel step (the fused exchange is
    // timed as whole steps below)
    bt_tuning split = prev;
    if (split.exchange_form == BT_EXCHANGE_FUSED) split.exchange_form = 1;
    bt_tuning best = split;
    double t[BT_AUTOTUNE_TRIALS];
    for (int q = 0; q < BT_AUTOTUNE_TRIALS; ++q) t[q] = -1.0;
    int rc = BT_OK;
    auto restore = [&]() {
        if (e == hipSuccess) e = state_copy(c, regions, save, true);
    };
    if (e == hipSuccess) {
        // membrane kernel: one or two wavefronts per tile
        double bm = 1e30;
        c->tune = split;
        for (int w = 1; w <= 2 && rc == BT_OK; ++w) {
            c->tune.mem_waves = w;
            rc = time_launches(c, 0, n_launches, &t[w - 1]);
            restore();
            if (rc == BT_OK && t[w - 1] < bm) { bm = t[w - 1]; best.mem_waves = w; }
        }
        c->tune = split;
        // env kernel: 1, 2, 3 or all ions per block
        const int cand[4] = {1, 2, 3, c->prm.n_ions};
        double be = 1e30;
        for (int q = 0; q < 4 && rc == BT_OK; ++q) {
            if (q < 3 && cand[q] >= c->prm.n_ions) continue;     // the same as all ions
            c->tune.env_ions_per_block = cand[q];
            rc = time_launches(c, 1, n_launches, &t[2 + q]);
            restore();
            if (rc == BT_OK && t[2 + q] < be) { be = t[2 + q]; best.env_ions_per_block = cand[q]; }
        }
        c->tune = split;
        // exchange: per ion or per square
        double bx = 1e30;
        for (int f = 0; f <= 1 && rc == BT_OK; ++f) {
            c->tune.exchange_form = f;
            rc = time_launches(c, 2, n_launches, &t[6 + f]);
            restore();
            if (rc == BT_OK && t[6 + f] < bx) { bx = t[6 + f]; best.exchange_form = f; }
        }
        // whole steps: replayed graphs on one stream, the persistent kernel,
        // replayed graphs on 2 and on 4 stream groups, with the variants
        // chosen above (the persistent kernel deals out its own work)
        if (rc == BT_OK && e == hipSuccess) {
            const int n_steps = 2 * std::max(n_launches, 4);
            double bs = 1e30;
            struct Whole { int slot, persistent, streams; bool fused; };
            const Whole cands[5] = {{8, 0, 1, false}, {9, 1, 1, false}, {10, 0, 2, false}, {11, 0, 4, false},
                                    {12, 0, 1, true}};
            bt_tuning fused_t = best;
            fused_t.exchange_form = BT_EXCHANGE_FUSED;
            fused_t.env_ions_per_block = c->prm.n_ions;
            for (const Whole& w : cands) {
                if (rc != BT_OK || e != hipSuccess) break;
                if (w.persistent && persistent_grid(c) < 1) continue;
                if (!streams_allowed(c, w.streams)) continue;
                if (w.fused && !fused_allowed(c, &fused_t)) continue;
                c->tune = w.fused ? fused_t : best;
                c->tune.persistent = w.persistent;
                c->tune.streams = w.streams;
                drop_graphs(c);
                if (!w.persistent) rc = capture_graphs(c);
                hipEvent_t e0, e1;
                if (rc == BT_OK) {
                    HIP_TRY(c, hipEventCreate(&e0));
                    HIP_TRY(c, hipEventCreate(&e1));
                    const int warm = 2 * std::max(1, (int)c->tune.graph_pairs);
                    rc = bt_step(c, warm, 1);                     // warm-up: every graph launched once
                    HIP_TRY(c, hipEventRecord(e0, c->stream));
                    if (rc == BT_OK) rc = bt_step(c, n_steps, 1);
                    HIP_TRY(c, hipEventRecord(e1, c->stream));
                    HIP_TRY(c, hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    unsigned fl = 0;
                    HIP_TRY(c, hipMemcpy(&fl, c->flags, sizeof(fl), hipMemcpyDeviceToHost));
                    if (fl & BT_FLAG_BARRIER) rc = set_err(c, BT_EHIP, "autotune: a grid barrier gave up waiting");
                    t[w.slot] = (double)ms / n_steps;
                    c->steps_done -= n_steps + warm;
                }
                restore();
                if (rc == BT_OK && t[w.slot] < bs) {
                    bs = t[w.slot];
                    if (w.fused) {
                        best.exchange_form = fused_t.exchange_form;
                        best.env_ions_per_block = fused_t.env_ions_per_block;
                    }
                    best.persistent = w.persistent;
                    best.streams = w.streams;
                }
            }
        }
        c->tune = prev;
    }
    (void)hipFree(save);
    if (rc != BT_OK) return rc;
    if (e != hipSuccess) return set_err(c, BT_EHIP, std::string("autotune state copy: ") + hipGetErrorString(e));
    c->tune = best;
    c->autotuned = true;
    drop_graphs(c);
    if (chosen) *chosen = best;
    if (trial_ms)
        for (int q = 0; q < n_trial_ms && q < BT_AUTOTUNE_TRIALS; ++q) trial_ms[q] = t[q];
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
    if (flags & BT_FLAG_BARRIER)
        return set_err(c, BT_EHIP, "a grid barrier of the persistent step kernel gave up waiting");
    if (flags) return set_err(c, BT_UNSTABLE, "NaN in cell concentrations or Vmem: simulation unstable");
    return BT_OK;
}

int bt_time_kernel(bt_ctx* c, int32_t kernel, int32_t n_launches, double* ms_out) {
    if (!c || !ms_out || n_launches < 1 || kernel < 0 || kernel > 4) return BT_EINVAL;
    if (kernel == 4 && !c->c_net) return set_err(c, BT_EINVAL, "no reaction network attached");
    return time_launches(c, kernel, n_launches, ms_out);
}

int bt_set_network(bt_ctx* c, const bt_network* net, double* c_net) {
    if (!c) return BT_EINVAL;
    if (net) {
        if (!c_net) return set_err(c, BT_EINVAL, "reaction network state buffer missing");
        if (net->n_species < 1 || net->n_species > BT_NET_MAX_SPECIES || net->n_rates < 0 ||
            net->n_rates > BT_NET_MAX_RATES)
            return set_err(c, BT_EINVAL, "reaction network size out of range");
        for (int j = 0; j < net->n_rates; ++j) {
            const bt_rate& r = net->rates[j];
            const int S = net->n_species;
            bool ok = (r.kind == BT_RATE_GROWTH || r.kind == BT_RATE_REACTION) &&
                      r.n_act >= 0 && r.n_act <= BT_NET_MAX_HILL && r.n_inh >= 0 && r.n_inh <= BT_NET_MAX_HILL &&
                      r.n_reac >= 0 && r.n_reac <= BT_NET_MAX_TERMS && r.n_prod >= 0 &&
                      r.n_prod <= BT_NET_MAX_TERMS && r.n_stoich >= 0 && r.n_stoich <= BT_NET_MAX_STOICH;
            if (ok && r.kind == BT_RATE_GROWTH) ok = r.species >= 0 && r.species < S;
            if (ok && r.kind == BT_RATE_REACTION && r.reversible) ok = r.n_reac >= 1 && r.n_prod >= 1;
            for (int q = 0; ok && q < r.n_act; ++q) ok = r.act[q] >= 0 && r.act[q] < S;
            for (int q = 0; ok && q < r.n_inh; ++q) ok = r.inh[q] >= 0 && r.inh[q] < S;
            for (int q = 0; ok && q < r.n_reac; ++q) ok = r.reac[q] >= 0 && r.reac[q] < S;
            for (int q = 0; ok && q < r.n_prod; ++q) ok = r.prod[q] >= 0 && r.prod[q] < S;
            for (int q = 0; ok && q < r.n_stoich; ++q) ok = r.stoich_species[q] >= 0 && r.stoich_species[q] < S;
            if (!ok) return set_err(c, BT_EINVAL, "reaction network rate " + std::to_string(j) + " is invalid");
        }
        // the kernel's form of the network: reciprocal constants, per-substance terms
        KNet k;
        memset(&k, 0, sizeof(k));
        k.n_species = net->n_species;
        k.n_rates = net->n_rates;
        for (int j = 0; j < net->n_rates; ++j) {
            const bt_rate& r = net->rates[j];
            KNetRate& q = k.rates[j];
            q.kind = r.kind; q.species = r.species; q.reversible = r.reversible;
            q.n_act = r.n_act; q.n_inh = r.n_inh; q.n_reac = r.n_reac; q.n_prod = r.n_prod;
            for (int h = 0; h < BT_NET_MAX_HILL; ++h) {
                q.act[h] = r.act[h]; q.inh[h] = r.inh[h];
                q.act_rkm[h] = (h < r.n_act) ? 1.0 / r.act_km[h] : 0.0;
                q.inh_rkm[h] = (h < r.n_inh) ? 1.0 / r.inh_km[h] : 0.0;
                q.act_n[h] = r.act_n[h]; q.inh_n[h] = r.inh_n[h];
            }
            for (int h = 0; h < BT_NET_MAX_TERMS; ++h) {
                q.reac[h] = r.reac[h]; q.prod[h] = r.prod[h];
                q.reac_coeff[h] = r.reac_coeff[h]; q.prod_coeff[h] = r.prod_coeff[h];
                q.reac_rkm[h] = (h < r.n_reac && !r.reversible) ? 1.0 / r.reac_km[h] : 0.0;
            }
            q.r_prod = r.r_prod; q.r_decay = r.r_decay; q.decay_max = r.decay_max;
            q.inv_gad = (r.kind == BT_RATE_GROWTH) ? 1.0 / r.gad_scale : 0.0;
            q.vmax = r.vmax;
            q.inv_keq = (r.kind == BT_RATE_REACTION && r.reversible) ? 1.0 / r.keq : 0.0;
        }
        int nt = 0;
        for (int sp = 0; sp < net->n_species; ++sp) {
            k.sub_ptr[sp] = nt;
            for (int j = 0; j < net->n_rates; ++j)
                for (int h = 0; h < net->rates[j].n_stoich; ++h)
                    if (net->rates[j].stoich_species[h] == sp) {
                        k.term_rate[nt] = j;
                        k.term_coeff[nt] = net->rates[j].stoich_coeff[h];
                        ++nt;
                    }
        }
        k.sub_ptr[net->n_species] = nt;
        if (!c->net) HIP_TRY(c, hipMalloc((void**)&c->net, sizeof(KNet)));
        HIP_TRY(c, hipMemcpyAsync(c->net, &k, sizeof(KNet), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->c_net = c_net;
        c->net_species = net->n_species;
        c->net_rates = net->n_rates;
    } else {
        c->c_net = nullptr;
    }
    // captured steps do not match the new step body any more, and neither
    // the persistent kernel nor stream groups have a network phase
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int g = 1; g < BT_MAX_STREAMS; ++g)
        if (c->gstream[g]) HIP_TRY(c, hipStreamSynchronize(c->gstream[g]));
    drop_graphs(c);
    c->pers_grid = -1;
    if (c->tune.persistent && persistent_grid(c) < 1) c->tune.persistent = 0;
    if (!streams_allowed(c, c->tune.streams)) c->tune.streams = 1;
    return BT_OK;
}

int bt_step_observed(bt_ctx* c, int32_t n_steps, int32_t use_graph, const bt_obs* obs) {
    if (!c || !obs) return BT_EINVAL;
    if (n_steps < 1) return set_err(c, BT_EINVAL, "an observed run needs at least one step");
    if (!obs->fluxes_mem || !obs->fluxes_gj || !obs->fluxes_env_x || !obs->fluxes_env_y ||
        !obs->rate_NaKATP || (c->prm.n_chan > 0 && !obs->chan_flux))
        return set_err(c, BT_EINVAL, "observation buffers missing");
    if (c->prm.i_na < 0)
        HIP_TRY(c, hipMemsetAsync(obs->rate_NaKATP, 0, (size_t)c->prm.n_mems * 8, c->stream));
    const size_t vbytes = (size_t)c->prm.n_cells * 8;
    if (!c->vm_prev) HIP_TRY(c, hipMalloc((void**)&c->vm_prev, vbytes));
    const size_t pbytes = (size_t)c->prm.n_squares * 8;
    if (c->buf.Phi_b && !c->phib_prev) HIP_TRY(c, hipMalloc((void**)&c->phib_prev, pbytes));
    if (c->phib_staged && n_steps > 1) {
        int rc = bt_step(c, 1, use_graph);       // the step the change fires in
        if (rc != BT_OK) return rc;
        --n_steps;
    }
    const int ahead = n_steps - 1;
    if (ahead > 0) {
        // keep the graph on the caller buffers: an even number of graph steps, then eager ones
        int rc = bt_step(c, ahead - (ahead & 1), use_graph);
        if (rc != BT_OK) return rc;
    }
    int rest = (ahead & 1) + 1;           // 1 or 2 eager steps, the last one observed
    if (rest == 2) {
        enqueue_step(c, 0);
        HIP_TRY(c, hipMemcpyAsync(c->vm_prev, c->vm_b, vbytes, hipMemcpyDeviceToDevice, c->stream));
        if (c->phib_prev)
            HIP_TRY(c, hipMemcpyAsync(c->phib_prev, c->buf.Phi_b, pbytes, hipMemcpyDeviceToDevice, c->stream));
        c->capture = obs;
        enqueue_step(c, 1);
        c->capture = nullptr;
    } else {
        HIP_TRY(c, hipMemcpyAsync(c->vm_prev, c->buf.vm_cell, vbytes, hipMemcpyDeviceToDevice, c->stream));
        if (c->phib_prev)
            HIP_TRY(c, hipMemcpyAsync(c->phib_prev, c->buf.Phi_b, pbytes, hipMemcpyDeviceToDevice, c->stream));
        c->capture = obs;
        enqueue_step(c, 0);
        c->capture = nullptr;
        int rc = copy_back(c);
        if (rc == BT_OK && c->phib_staged) rc = apply_staged_boundary(c);
        if (rc != BT_OK) return rc;
    }
    enqueue_step(c, 0, 3);
    HIP_TRY(c, hipGetLastError());
    c->vm_prev_ok = true;
    c->steps_done += rest;
    return BT_OK;
}

static int launch_gemm(bt_ctx* c, const double* A, const double* B, double* C, int m, int n, int k) {
    const int tiles = (m / 16) * (n / 16);
    hipLaunchKernelGGL(k_gemm_f64, dim3(tiles), dim3(256), 0, c->stream, A, B, C, m, n, k);
    HIP_TRY(c, hipGetLastError());
    return BT_OK;
}

int bt_gemm_f64(bt_ctx* c, const double* A, const double* B, double* C, int32_t m, int32_t n, int32_t k) {
    if (!c || !A || !B || !C || m < 16 || n < 16 || k < 16 || (m | n | k) % 16)
        return set_err(c, BT_EINVAL, "gemm sizes must be positive multiples of 16");
    return launch_gemm(c, A, B, C, m, n, k);
}

int bt_fp32_math(const float* x, const float* y, float* out, int32_t n, int32_t fn, void* stream) {
    if (!x || !out || n < 0 || fn < 0 || fn > 2 || (fn == 2 && !y)) return BT_EINVAL;
    if (n == 0) return BT_OK;
    hipLaunchKernelGGL(k_fp32_math, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, y, out, n, fn);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_fp64_math(const double* x, const double* y, double* out, int32_t n, int32_t fn, void* stream) {
    if (!x || !out || n < 0 || fn < 0 || fn > 2 || (fn == 2 && !y)) return BT_EINVAL;
    if (n == 0) return BT_OK;
    hipLaunchKernelGGL(k_fp64_math, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, y, out, n, fn);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

// The observables workspace: per-square buffers for the batch, DST bases per
// interior size, per-cluster eigenvalue tables and padded matrices, and the
// product tables of the four DST stages. A cluster's Dirichlet Poisson solve
// by DST-I is X = S1 ((S1 B S2) .* invlam) S2 (S the symmetric DST-I matrix
// of the interior's size, invlam the inverse eigenvalues with the transform's
// normalisation and delta^2); every cluster's two solves (stream function and
// potential) are products of one grouped launch per stage.
static int obs_workspace(bt_ctx* c) {
    if (c->obs_ready) return BT_OK;
    auto dalloc = [&](void** ptr, size_t bytes) -> int {
        HIP_TRY(c, hipMalloc(ptr, bytes < 256 ? 256 : bytes));
        c->obs_alloc.push_back(*ptr);
        return BT_OK;
    };
    const size_t ns = (size_t)c->kp.n_sq;
    int rc;
    if ((rc = dalloc((void**)&c->jxo, ns * 8)) || (rc = dalloc((void**)&c->jyo, ns * 8))) return rc;
    // DST-I basis S[j][k] = sin(pi (j+1)(k+1) / (n+1)), zero in the padding
    std::vector<std::pair<int, double*>> bases;
    auto basis = [&](int n, double** out) -> int {
        for (auto& b : bases)
            if (b.first == n) { *out = b.second; return BT_OK; }
        const int np = (n + 15) / 16 * 16;
        std::vector<double> h((size_t)np * np, 0.0);
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) h[(size_t)j * np + k] = sin(M_PI * (j + 1) * (k + 1) / (n + 1));
        int r = dalloc((void**)out, h.size() * 8);
        if (r != BT_OK) return r;
        HIP_TRY(c, hipMemcpy(*out, h.data(), h.size() * 8, hipMemcpyHostToDevice));
        bases.push_back({n, *out});
        return BT_OK;
    };
    const size_t nq = c->kcl_host.size();
    std::vector<ObsQ> qs(nq);
    std::vector<double*> S1(nq), S2(nq);
    long long off = 0;
    for (size_t q = 0; q < nq; ++q) {
        const KCluster& K = c->kcl_host[q];
        ObsQ& g = qs[q];
        g.n1 = K.ny - 2; g.n2 = K.nx - 2;
        g.n1p = (g.n1 + 15) / 16 * 16; g.n2p = (g.n2 + 15) / 16 * 16;
        g.off = off;
        off += 2LL * g.n1p * g.n2p;
        if ((rc = basis(g.n1, &S1[q])) || (rc = basis(g.n2, &S2[q]))) return rc;
    }
    const size_t wbytes = (size_t)off * 8;
    if ((rc = dalloc((void**)&c->Bw, wbytes)) || (rc = dalloc((void**)&c->T1w, wbytes)) ||
        (rc = dalloc((void**)&c->T2w, wbytes)))
        return rc;
    // inverse eigenvalues of the 1-D second differences 2 cos(pi (j+1)/(n+1)) - 2,
    // with the transform's normalisation and delta^2, per cluster (zero in the padding)
    std::vector<double*> invlam(nq);
    for (size_t q = 0; q < nq; ++q) {
        const KCluster& K = c->kcl_host[q];
        const ObsQ& g = qs[q];
        std::vector<double> hl((size_t)g.n1p * g.n2p, 0.0);
        const double dl2 = K.delta * K.delta;
        const double scale = 4.0 / ((double)(g.n1 + 1) * (double)(g.n2 + 1));
        for (int j = 0; j < g.n1; ++j)
            for (int k = 0; k < g.n2; ++k) {
                const double lam = (2.0 * cos(M_PI * (j + 1) / (g.n1 + 1)) - 2.0) +
                                   (2.0 * cos(M_PI * (k + 1) / (g.n2 + 1)) - 2.0);
                hl[(size_t)j * g.n2p + k] = dl2 * scale / lam;
            }
        if ((rc = dalloc((void**)&invlam[q], hl.size() * 8))) return rc;
        HIP_TRY(c, hipMemcpy(invlam[q], hl.data(), hl.size() * 8, hipMemcpyHostToDevice));
    }
    // the four stages: T1 = S1 B; T2 = (T1 S2) .* invlam; T1 = S1 T2; T2 = T1 S2
    const int np = (int)(2 * nq);
    std::vector<GemmProb> probs((size_t)4 * np);
    std::vector<int32_t> bmap;
    for (size_t q = 0; q < nq; ++q)
        for (int h = 0; h < 2; ++h) {
            const ObsQ& g = qs[q];
            const size_t o = (size_t)g.off + (size_t)h * g.n1p * g.n2p;
            const int p = (int)(2 * q + h);
            const int tile0 = (int)bmap.size();
            const int tiles = (g.n1p / 16) * (g.n2p / 16);
            for (int t = 0; t < tiles; ++t) bmap.push_back(p);
            double* B = c->Bw + o; double* T1 = c->T1w + o; double* T2 = c->T2w + o;
            probs[(size_t)0 * np + p] = GemmProb{S1[q], B, T1, nullptr, g.n1p, g.n2p, g.n1p, tile0};
            probs[(size_t)1 * np + p] = GemmProb{T1, S2[q], T2, invlam[q], g.n1p, g.n2p, g.n2p, tile0};
            probs[(size_t)2 * np + p] = GemmProb{S1[q], T2, T1, nullptr, g.n1p, g.n2p, g.n1p, tile0};
            probs[(size_t)3 * np + p] = GemmProb{T1, S2[q], T2, nullptr, g.n1p, g.n2p, g.n2p, tile0};
        }
    if ((rc = dalloc((void**)&c->gemm_probs, probs.size() * sizeof(GemmProb))) ||
        (rc = dalloc((void**)&c->block_prob, bmap.size() * 4)) || (rc = dalloc((void**)&c->obs_q, nq * sizeof(ObsQ))))
        return rc;
    HIP_TRY(c, hipMemcpy(c->gemm_probs, probs.data(), probs.size() * sizeof(GemmProb), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->block_prob, bmap.data(), bmap.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->obs_q, qs.data(), nq * sizeof(ObsQ), hipMemcpyHostToDevice));
    c->n_pad = off / 2;
    c->n_prob = np;
    c->n_gemm_blocks = (int)bmap.size();
    c->obs_ready = true;
    return BT_OK;
}

int bt_observables(bt_ctx* c, const bt_obs* o) {
    if (!c || !o) return BT_EINVAL;
    const void* need[] = {o->fluxes_mem, o->fluxes_gj, o->fluxes_env_x, o->fluxes_env_y, o->Jn, o->I_mem,
                          o->E_gj_x, o->E_gj_y, o->J_cell_x, o->J_cell_y, o->vm_ave, o->J_env_x, o->J_env_y,
                          o->Jtx, o->Jty, o->B_field};
    for (const void* q : need)
        if (!q) return set_err(c, BT_EINVAL, "observable buffer missing");
    if (c->prm.n_chan > 0 && !o->chan_flux) return set_err(c, BT_EINVAL, "chan_flux missing");
    if (!c->buf.mem_nx || !c->buf.mem_ny || !c->buf.cell_sa)
        return set_err(c, BT_EINVAL, "observables need mem_nx, mem_ny and cell_sa");
    int rc = obs_workspace(c);
    if (rc != BT_OK) return rc;
    const bt_buffers& B = c->buf;
    ObsMemArgs a{B.mem_ptr, B.nn_mem, B.mem_sa, B.mem_nx, B.mem_ny, B.cell_sa, B.vm_cell,
                 c->vm_prev_ok ? c->vm_prev : B.vm_cell, B.mem_cell, B.tile_ptr,
                 o->fluxes_mem, o->fluxes_gj, o->chan_flux, o->Jn, o->I_mem, o->E_gj_x, o->E_gj_y,
                 o->J_cell_x, o->J_cell_y, o->vm_ave, c->prm.smooth_cells,
                 B.Phi_b, (c->vm_prev_ok && c->phib_prev) ? c->phib_prev : B.Phi_b, B.mem_sq, o->vm};
    hipLaunchKernelGGL(k_obs_mem, dim3(c->n_tiles), dim3(TMB), 0, c->stream, a, c->kp);
    hipLaunchKernelGGL(k_obs_env_rhs, dim3(grid_blocks(c->kp.n_sq)), dim3(BLOCK), 0, c->stream,
                       (const double*)o->fluxes_env_x, (const double*)o->fluxes_env_y, c->jxo, c->jyo, c->kp);
    // the Helmholtz-Hodge decomposition of each cluster's env current on its
    // own grid: both right-hand sides, the four DST stages of both solves of
    // every cluster (one grouped launch each), then the outputs
    const int nq = (int)c->kcl_host.size();
    hipLaunchKernelGGL(k_obs_rhs, dim3((unsigned)((c->n_pad + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, c->stream,
                       (const double*)c->jxo, (const double*)c->jyo, c->Bw, (const ObsQ*)c->obs_q, nq, c->n_pad,
                       (const KCluster*)c->kcl);
    for (int st = 0; st < 4; ++st)
        hipLaunchKernelGGL(k_gemm_grouped, dim3(c->n_gemm_blocks), dim3(256), 0, c->stream,
                           (const GemmProb*)(c->gemm_probs + (size_t)st * c->n_prob), (const int32_t*)c->block_prob);
    hipLaunchKernelGGL(k_obs_hh_out, dim3(grid_blocks(c->kp.n_sq)), dim3(BLOCK), 0, c->stream, (const double*)c->T2w,
                       (const ObsQ*)c->obs_q, (const int32_t*)B.sq_cluster, (const KCluster*)c->kcl, o->J_env_x,
                       o->J_env_y, o->Jtx, o->Jty, o->B_field, c->prm.mu, c->kp.n_sq,
                       c->any_phib ? B.Phi_b : nullptr);
    HIP_TRY(c, hipGetLastError());
    return BT_OK;
}

int bt_residual(bt_ctx* c, const double* c_prev, const int32_t* cell_cluster, double* r_out) {
    if (!c || !c_prev || !cell_cluster || !r_out) return BT_EINVAL;
    const size_t nq = c->clusters.size();
    HIP_TRY(c, hipMemsetAsync(r_out, 0, nq * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_residual, dim3(grid_blocks(c->kp.n_cells)), dim3(BLOCK), 0, c->stream,
                       (const double*)c->buf.c_cell, c_prev, cell_cluster, (unsigned long long*)r_out,
                       c->kp.n_cells, c->kp.n_ions);
    HIP_TRY(c, hipGetLastError());
    return BT_OK;
}

}  // extern "C"
