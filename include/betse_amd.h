/*
 * betse_amd.h -- C ABI of the MI355X (gfx950) electrodiffusion step.
 *
 * The reference has no FFI: its seam is the solver dispatch in
 * Simulator.run_sim_core (betse/science/sim.py:1064-1102), which calls
 * Simulator._run_sim_core_loop(phase, time_steps, time_steps_sampled,
 * anim_cells) (sim.py:1132-1138). Each entry point below replaces a part of
 * that loop body; the Python host (betse_amd/sim.py) binds them with ctypes.
 *
 * All pointers are device pointers owned by the caller (PyTorch tensors in the
 * Python host), except the workspace the library allocates in bt_create and
 * frees in bt_destroy. Every call is asynchronous on the stream given to
 * bt_create; bt_sync waits for it. Return codes: BT_OK (0), BT_UNSTABLE (>0,
 * a NaN was produced: BetseSimUnstableException in the reference,
 * sim_toolbox.py:332-344 and 475-503), BT_EINVAL / BT_EHIP (<0, see
 * bt_last_error).
 */
#ifndef BETSE_AMD_H
#define BETSE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BT_OK        0
#define BT_UNSTABLE  1
#define BT_EINVAL   -1
#define BT_EHIP     -2

#define BT_MAX_IONS  8
#define BT_MAX_CHAN  4
#define BT_TILE_MEMBRANES 64   /* at most this many membranes per cell tile (whole cells, one wavefront) */
#define BT_TILE_CELLS     32   /* at most this many cells per cell tile */

/* Voltage-gated channel models (betse/science/channels/vg_na.py, vg_k.py, vg_ca.py). */
#define BT_CHAN_NAV1P2 0
#define BT_CHAN_KV1P1  1
#define BT_CHAN_CAV1P2 2

/* Flux arithmetic of the membrane and environment flux laws.
 * Cell/environment state, charge sums and the charge->voltage step are FP64
 * in both modes. */
#define BT_FLUX_FP64 0
#define BT_FLUX_FP32 1

typedef struct bt_ctx bt_ctx;

/* One cluster of a batch: its environment grid and the constants that differ
 * between independent clusters (host memory, copied by bt_create). A single
 * simulation is a batch of one. Cells, membranes and squares of all clusters
 * are concatenated in the bt_buffers arrays; indices are global. */
typedef struct {
    int32_t grid_ny, grid_nx;           /* this cluster's environment grid     */
    int32_t sq_off;                     /* first square of the grid           */
    int32_t reserved;
    double delta;                       /* grid spacing (cells.py:1607)        */
    double mean_sa;                     /* memSa_per_envSquare[map_mem2ecm].mean() (ion_current.py:93-95) */
    double ko_env;                      /* frozen per phase (sim.py:974, Q13) */
    double alpha_NaK;                   /* NaK-ATPase rate (sim_config.yaml:1030) */
} bt_cluster;

/* Sizes, constants and switches; the names follow betse/science/parameters.py. */
typedef struct {
    int32_t n_cells, n_mems, n_squares, reserved0, n_ions;
    int32_t i_na, i_k, i_ca;            /* ion slots, -1 if the ion is absent */
    int32_t flux_mode;                  /* BT_FLUX_FP64 / BT_FLUX_FP32 */
    int32_t n_chan;                     /* voltage-gated channels on every membrane (0..BT_MAX_CHAN) */
    int32_t chan_kind[BT_MAX_CHAN];     /* BT_CHAN_* */
    int32_t chan_ion[BT_MAX_CHAN];      /* ion slot the channel conducts; one channel per ion */
    double chan_max_dm[BT_MAX_CHAN];    /* chan.maxDm (networks.py:6557) */
    double chan_dt;                     /* p.dt * time_unit (ms), channelsabc.py:54 */
    double z[BT_MAX_IONS];              /* valence (parameters.py:1082-1088) */
    double d_gj[BT_MAX_IONS];           /* D_gj per ion (sim.py:570-574) */
    double c_bound[BT_MAX_IONS];        /* c_env_bound (sim.py:608, 800-806) */
    double dt, T, T_gj;                 /* sim.T for membrane/pump/env, p.T for GJ (Q7) */
    double F, R, q, kb, cm, tm;
    double gj_len, gj_surface;
    double NaK_block, deltaGATP, cATP, cADP, cPi;
    double KmNK_Na, KmNK_K, KmNK_ATP;
    double alpha_Ca, KmCa_Ca, KmCa_ATP;
    double cell_height;
    double eo, er;
    double cell_radius, true_cell_size;
    double gauss_w[5];                  /* scipy gaussian_filter1d taps, centre first */
    double smooth_cells;                /* 1/'sharpness cell' (parameters.py:1146), Jn smoothing */
    double mu;                          /* magnetic permeability (parameters.py:1111), B_field */
} bt_params;

/* Device buffers. Layouts: per-ion fields are [n_ions][n], each cluster's grid
 * is row-major [grid_ny][grid_nx] (rows = y = axis 0) at its sq_off, indices
 * are int32 and global over the batch. */
typedef struct {
    /* geometry (read only) */
    const int32_t* mem_ptr;     /* [n_cells+1] CSR of each cell's membranes       */
    const int32_t* mem_cell;    /* [n_mems] mem_to_cells                          */
    const int32_t* nn_mem;      /* [n_mems] nn_i (self for boundary membranes)    */
    const int32_t* nn_cell;     /* [n_mems] mem_to_cells[nn_i]                    */
    const int32_t* mem_sq;      /* [n_mems] map_mem2ecm                           */
    const int32_t* sq_ptr;      /* [ny*nx+1] grid-side CSR of M_divmap_mem2ecm    */
    const int32_t* sq_mems;     /* [n_mems] membranes sorted by square            */
    const int32_t* mem_pos;     /* [n_mems] position of each membrane in sq_mems  */
    const int32_t* tile_ptr;    /* [n_tiles+1] first cell of each tile of whole cells
                                   of one cluster holding <= BT_TILE_MEMBRANES membranes
                                   and <= BT_TILE_CELLS cells */
    const int32_t* tile_cluster;/* [n_tiles] cluster of each tile                 */
    const int32_t* sq_cluster;  /* [n_squares] cluster of each square             */
    const double* mem_sa;       /* [n_mems]                                       */
    const double* gjopen;       /* [n_mems] gj_block*gj_default_weights           */
    const double* cell_vol;     /* [n_cells]                                      */
    const double* diviterm;     /* [n_cells]                                      */
    const double* Dm;           /* [n_ions][n_mems] Dm_cells                      */
    const double* D_env;        /* [n_ions][ny*nx] D_env*TJ_modulator             */
    const uint8_t* sq_has_mem;  /* [ny*nx] 1 where some membrane maps            */
    /* state (read/write, in place between calls) */
    double* c_cell;             /* [n_ions][n_cells] cc_cells                     */
    double* c_lag;              /* [n_ions][n_cells] cc_at_mem per cell (Q1)      */
    double* vm_cell;            /* [n_cells] (1/cm) rho_surf of each cell: the Vmem
                                   of its membranes before - Phi_b[map_mem2ecm]  */
    double* c_env;              /* [n_ions][ny*nx] cc_env                         */
    double* E_x;                /* [ny*nx] E_env_x                                */
    double* E_y;                /* [ny*nx] E_env_y                                */
    double* v_env;              /* [ny*nx] v_env                                  */
    /* geometry used by the sampled-step observables only */
    const double* mem_nx;       /* [n_mems] membrane normal x (mem_vects_flat[:,2]) */
    const double* mem_ny;       /* [n_mems] membrane normal y (mem_vects_flat[:,3]) */
    const double* cell_sa;      /* [n_cells]                                      */
    /* voltage-gated channels (may be NULL when n_chan == 0) */
    double* gates;              /* [n_chan][2][n_mems] m, h gates (in place)      */
    const double* chan_mod;     /* [n_chan][n_cells] channel modulator per cell   */
    /* applied boundary voltage (bound_V events; ion_current.py:79-90): Phi_b
       = lapENVinv . (-div_Jb) per square, subtracted from each membrane's Vmem
       and added to v_env; NULL or all zeros: none. Not with channels. State:
       bt_stage_boundary writes a staged change into it. */
    double* Phi_b;              /* [ny*nx]                                        */
    int32_t n_tiles;
    int32_t tile_m;             /* must equal BT_TILE_MEMBRANES                   */
} bt_buffers;

/* Per-cell reaction network of intracellular substances
 * (MasterOfNetworks.run_loop, betse/science/chemistry/networks.py:2805-2982).
 * Every rate is either a substance's growth/decay term (write_growth_and_decay,
 * networks.py:1187-1305) or a reaction (write_reactions, networks.py:1312-1570),
 * each scaled by Hill activators and inhibitors (get_influencers,
 * networks.py:5084-5600). The substances stay inside their cells (no membrane
 * or gap-junction transport), so a step is a per-cell explicit Euler update
 * c += (reaction_matrix . rates) dt followed by no_negs. */
#define BT_NET_MAX_SPECIES 16
#define BT_NET_MAX_RATES   32
#define BT_NET_MAX_HILL     4
#define BT_NET_MAX_TERMS    4
#define BT_NET_MAX_STOICH   8
#define BT_RATE_GROWTH      0
#define BT_RATE_REACTION    1

typedef struct {
    int32_t kind;                       /* BT_RATE_GROWTH or BT_RATE_REACTION */
    int32_t species;                    /* growth: its substance               */
    int32_t reversible;                 /* reaction with delta_Go: vmax*alpha*(1 - Q/Keq) */
    int32_t n_act, n_inh;               /* Hill activators and inhibitors      */
    int32_t n_reac, n_prod;             /* reaction terms                      */
    int32_t n_stoich;                   /* nonzero entries of this rate's reaction_matrix column */
    int32_t act[BT_NET_MAX_HILL];
    int32_t inh[BT_NET_MAX_HILL];
    int32_t reac[BT_NET_MAX_TERMS];
    int32_t prod[BT_NET_MAX_TERMS];
    int32_t stoich_species[BT_NET_MAX_STOICH];
    double act_km[BT_NET_MAX_HILL], act_n[BT_NET_MAX_HILL];
    double inh_km[BT_NET_MAX_HILL], inh_n[BT_NET_MAX_HILL];
    double reac_coeff[BT_NET_MAX_TERMS], reac_km[BT_NET_MAX_TERMS];
    double prod_coeff[BT_NET_MAX_TERMS], prod_km[BT_NET_MAX_TERMS];
    double stoich_coeff[BT_NET_MAX_STOICH];
    double r_prod, r_decay, decay_max, gad_scale;   /* growth/decay (time factor folded in) */
    double vmax, keq;                               /* reaction; keq = exp(-delta_Go/(R T)) */
} bt_rate;

typedef struct {
    int32_t n_species, n_rates;
    bt_rate rates[BT_NET_MAX_RATES];
} bt_network;

/* Create a context: validates sizes, tiles and the grid-side CSR, copies the
 * parameter block, allocates the workspace (second cell-state and env-state
 * buffers, post-diffusion env buffer, per-membrane exchange fluxes, per-cell
 * gating rates, surface-charge field, the list of squares with membranes,
 * status words) on the current device. It reads tile_ptr, mem_ptr,
 * tile_cluster, sq_ptr, sq_has_mem, Dm, gjopen and D_env with blocking
 * copies, so their contents must be complete when it is called. Dm, gjopen
 * and D_env are read once here (a cluster whose values are uniform gets them
 * as constants); they do not change within a phase on this path, and a caller
 * that changes them creates a new context. Between calls the caller's state buffers hold the
 * state (the workspace copies are internal).
 * `stream` is a hipStream_t (may be NULL for the default stream; graphs need a
 * non-default stream). Replaces the state set-up of init_core/init_dynamics
 * (sim.py:452-1012) on the device side. */
int bt_create(bt_ctx** out, const bt_params* params, const bt_cluster* clusters, int32_t n_clusters,
              const bt_buffers* buffers, void* stream);

/* Check the geometry arrays of a batch on the host before any kernel gathers
 * through them (no counterpart in the reference, whose NumPy indexing raises
 * IndexError on an out-of-range index). `host` holds HOST pointers to
 * mem_ptr, mem_cell, nn_mem, nn_cell, mem_sq, sq_ptr, sq_mems, mem_pos,
 * tile_ptr, tile_cluster and sq_cluster (the other fields are not read) and
 * n_tiles, tile_m. Checks: the cell CSR and the tiles cover every cell and
 * membrane within the tile limits; mem_cell is the CSR owner of each
 * membrane; nn_mem is a membrane of the same cluster and nn_cell its cell;
 * mem_sq is a square of the membrane's cluster; sq_cluster matches the
 * cluster grids; sq_ptr/sq_mems list every membrane once under its square and
 * mem_pos is the inverse permutation. Returns BT_OK or BT_EINVAL (message in
 * bt_last_error(NULL)). bt_create runs the same checks on copies of the
 * device arrays. Touches no device. */
int bt_validate(const bt_params* params, const bt_cluster* clusters, int32_t n_clusters,
                const bt_buffers* host);

/* Release the workspace, graphs and the context. */
int bt_destroy(bt_ctx* ctx);

/* Launch variants of the step kernels (no counterpart in the reference: the
 * NumPy loop has no launches). bt_create picks them from the problem size;
 * bt_set_tuning forces them (captured step graphs are rebuilt) and
 * bt_autotune times every variant on the context's own problem and keeps the
 * fastest of each kernel. The step's arithmetic is the same in every variant:
 * only the division of the work between wavefronts and blocks changes. */
typedef struct {
    int32_t mem_waves;          /* k_mem_cell wavefronts per cell tile: 1, or 2 (each a contiguous block of ions) */
    int32_t env_ions_per_block; /* ions per k_env_field block, 1..n_ions (n_ions: one block per 16x16 tile) */
    int32_t exchange_form;      /* k_env_exchange: 0 one wavefront per ion over 64 squares, 1 one thread per square;
                                   BT_EXCHANGE_FUSED: no exchange kernel, k_mem_cell runs first and k_env_field
                                   adds each square's membrane fluxes to its diffused value (needs no channels, no
                                   Ca pump and env_ions_per_block = n_ions) */
    int32_t graph_pairs;        /* pairs of steps per replayed step graph, 1..64 */
    int32_t persistent;         /* 1: bt_step runs its pairs of steps in cooperative launches of one persistent
                                   kernel (grid barriers between a step's phases; 5 or 6 ions, no reaction
                                   network); 0: replayed step graphs */
    int32_t streams;            /* 1..BT_MAX_STREAMS: with replayed graphs, the clusters of a batch are split
                                   into this many groups of about equal cells, each stepped by its own graphs
                                   on its own stream, so that one group's kernels run in the ramps and tails of
                                   the others' (clusters do not interact within a step); more than 1 needs at
                                   least that many clusters, no channels and no reaction network */
} bt_tuning;

#define BT_MAX_STREAMS 4
#define BT_EXCHANGE_FUSED 2

/* Events that change the membranes' or the environment's diffusivities
 * (TissueHandler.fire_events -> makeAllChanges, tishandler.py:709-724 and
 * 1321-1332: Dm_cells = Dm_scheduled + Dm_base; gap-junction blocks through
 * gjopen; D_env): after the caller has written new values into the Dm, gjopen,
 * D_env (or Phi_b) buffers bound at bt_create, this re-derives the per-cluster
 * constants bt_create took from them (values that are one over a cluster, ions
 * without membrane flux, clusters with an applied voltage), so the next steps
 * use the new values; captured graphs stay valid. BT_EINVAL for a nonzero
 * Phi_b with voltage-gated channels. Synchronous. */
int bt_refresh_params(bt_ctx* ctx);

/* The applied boundary voltage event (TissueEventVoltage.fire,
 * tisevevolt.py:75-89, sets sim.bound_V at the start of a step; get_current,
 * ion_current.py:84-107, turns it into Phi_b at that step's end): Phi_b_next
 * is the new potential per square (host memory, n_squares values: the harmonic
 * extension of bound_V that get_current solves for, zero where a cluster has
 * none). The next step still runs with the Phi_b in effect; from its end on,
 * Phi_b_next is copied into the Phi_b buffer bound at bt_create and the
 * per-cluster constants follow (as bt_refresh_params): Vmem per membrane is
 * its cell's less Phi_b at its square, v_env adds Phi_b, and the observables of
 * that step take E_gj from the Vmem before it and the rest from after it, as
 * the reference does. BT_EINVAL without a Phi_b buffer, for a non-finite value,
 * or for a nonzero value with voltage-gated channels (their gating rates are
 * per cell). Synchronous; bt_autotune refuses while a change is staged. */
int bt_stage_boundary(bt_ctx* ctx, const double* Phi_b_next);

/* Capture and instantiate the step graphs now (bt_step otherwise does it on
 * its first call that replays graphs), so that no later bt_step pays for it.
 * Newly captured graphs are also replayed once from a saved copy of the state,
 * which is then put back (bit for bit), so that their first launch's one-time
 * costs fall here too. Allocates the copy for the duration of the call. */
int bt_capture_graphs(bt_ctx* ctx);

int bt_get_tuning(const bt_ctx* ctx, bt_tuning* out);
int bt_set_tuning(bt_ctx* ctx, const bt_tuning* tuning);

/* Time each kernel variant over n_launches launches (HIP events on the
 * context stream) and keep the fastest per kernel, then time whole steps
 * (2 * max(n_launches, 4)) replayed from graphs on one stream, in the
 * persistent kernel where that exists, and replayed from graphs on 2 and on 4
 * streams where the problem allows it; the chosen variants go to *chosen (may
 * be NULL) and the average ms of each trial to trial_ms (may be NULL;
 * n_trial_ms entries, unused ones -1) in the order: mem_waves 1, 2;
 * env_ions_per_block 1, 2, 3, n_ions; exchange_form 0, 1 (ms per launch);
 * persistent 0, 1; streams 2, 4; exchange_form BT_EXCHANGE_FUSED with all ions
 * per env block (ms per step).
 * All the state a step modifies in place is saved before and restored after
 * each trial, so a phase continues from where it was; the state must be
 * stable (no NaN flagged). Allocates a temporary copy of that state.
 * Synchronous. */
int bt_autotune(bt_ctx* ctx, int32_t n_launches, bt_tuning* chosen, double* trial_ms, int32_t n_trial_ms);

/* update_V's voltage part for the current state (sim.py:2007-2083 with
 * ion_current.py:13-114): rho_cells -> vm_cell, rho_env -> Gaussian -> v_env
 * -> E_env. Called once before a phase's loop (sim.py:1041). */
int bt_update_v(bt_ctx* ctx);

/* Enqueue n_steps iterations of the _run_sim_core_loop body (sim.py:1169-1366):
 * NaK pump, per-ion GHK + gap-junction fluxes, env Nernst-Planck update,
 * Ca pump, voltage-gated channels (networks.py:3115-3215), cell/env
 * concentration update, update_V, NaN checks. Steps are
 * replayed from a captured hipGraph when use_graph != 0. Asynchronous. */
int bt_step(bt_ctx* ctx, int32_t n_steps, int32_t use_graph);

/* Attach a reaction network to the step (NULL net detaches it). c_net is the
 * caller's device buffer [n_species][n_cells] of substance concentrations
 * (mM), updated in place by every following step after the channels and
 * before the ion concentration update, as in sim.py:1286-1300. A substance
 * that becomes NaN halts the run like a NaN ion concentration (no_negs on
 * the substances, networks.py:2939 -> sim_toolbox.py:475-503): bt_sync then
 * returns BT_UNSTABLE with that step's index. The network is copied;
 * captured step graphs are rebuilt. */
int bt_set_network(bt_ctx* ctx, const bt_network* net, double* c_net);

/* Observables of one step (get_current, ion_current.py:13-109; update_gj
 * E_gj, sim.py:2166-2172; update_V vm_ave, sim.py:2038). The reference
 * evaluates them every step; here they are produced for the last step of a
 * bt_step_observed call. All pointers are caller-owned device buffers; in a
 * batch the per-square fields are concatenated like c_env (each cluster's
 * grid at its sq_off). */
typedef struct {
    /* raw fluxes of the observed step, written by bt_step_observed */
    double* fluxes_mem;    /* [n_ions][n_mems] fluxes_mem (pumps + GHK)          */
    double* fluxes_gj;     /* [n_ions][n_mems] fluxes_gj                         */
    double* fluxes_env_x;  /* [n_ions][ny*nx] fluxes_env_x                       */
    double* fluxes_env_y;  /* [n_ions][ny*nx] fluxes_env_y                       */
    double* chan_flux;     /* [n_chan][n_mems] channel fluxes (NULL if n_chan==0) */
    double* rate_NaKATP;   /* [n_mems] -f_Na of the NaK pump                     */
    /* derived by bt_observables */
    double* Jn;            /* [n_mems] smoothed transmembrane current density   */
    double* I_mem;         /* [n_mems]                                           */
    double* E_gj_x;        /* [n_mems]                                           */
    double* E_gj_y;        /* [n_mems]                                           */
    double* J_cell_x;      /* [n_cells]                                          */
    double* J_cell_y;      /* [n_cells]                                          */
    double* vm_ave;        /* [n_cells]                                          */
    double* J_env_x;       /* [ny*nx] divergence-free env current (HH_Decomp)    */
    double* J_env_y;       /* [ny*nx]                                            */
    double* Jtx;           /* [ny*nx] J_env + curl-free part                     */
    double* Jty;           /* [ny*nx]                                            */
    double* B_field;       /* [ny*nx] mu * stream function                       */
    double* vm;            /* [n_mems] Vmem per membrane (update_V's vm, what
                              write2storage stores), or NULL                     */
} bt_obs;

/* bt_step whose last step also writes the raw fluxes into obs. */
int bt_step_observed(bt_ctx* ctx, int32_t n_steps, int32_t use_graph, const bt_obs* obs);

/* Derive the observables from the fluxes captured in obs. The two Poisson
 * solves of HH_Decomp (sim_toolbox.py:1236-1290, dense lapENVinv in the
 * reference) run as DST-I transforms written as FP64 MFMA matrix products,
 * one pair per cluster of a batch (each on its own grid). The DST bases and
 * eigenvalue tables are allocated and built on the first call. */
int bt_observables(bt_ctx* ctx, const bt_obs* obs);

/* C = A*B for row-major FP64 matrices whose sizes are multiples of 16, on the
 * context stream with the MFMA kernel the observables use (exposed for tests). */
int bt_gemm_f64(bt_ctx* ctx, const double* A, const double* B, double* C,
                int32_t m, int32_t n, int32_t k);

/* The FP32-flux path's elementary functions on n device floats, asynchronous
 * on `stream` (exposed for tests; no counterpart in the reference, whose laws
 * are FP64): fn 0 out = exp(x), 1 out = expm1(x) for |x| <= 1/2, 2 out = x / y. */
int bt_fp32_math(const float* x, const float* y, float* out, int32_t n, int32_t fn, void* stream);

/* The FP64 path's own elementary functions on n device doubles, asynchronous
 * on `stream` (exposed for tests; the reference calls NumPy's exp, expm1 and
 * division, sim_toolbox.py:18-122): fn 0 out = exp(x), 1 out = expm1(x) for
 * |x| <= 1/2, 2 out = x / y. */
int bt_fp64_math(const double* x, const double* y, double* out, int32_t n, int32_t fn, void* stream);

/* ------------------------------------------------------------------------
 * Warm-start graph network (no counterpart in the reference: it predicts the
 * state the init phase reaches, from which the plain stepper then runs to its
 * own residual tolerance). FP32 device buffers, asynchronous on `stream`.
 * ------------------------------------------------------------------------ */

/* C(m x n) = op(A) op(B) (beta=0) or C += op(A) op(B) (beta=1), where
 * A(i,k) = A[i*sam + k*sak] and B(k,j) = B[k*sbk + j*sbn]; operands are rounded
 * to BF16 and multiplied on v_mfma_f32_16x16x16bf16_1k with FP32 accumulation.
 * K is split over ksplit blocks whose partials go to ws (ksplit*m*n floats);
 * ws is also needed when beta=1. */
int bt_gnn_gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn,
                float* C, int64_t ldc, int32_t m, int32_t n, int32_t k, int32_t beta,
                float* ws, int32_t ksplit, void* stream);
/* Z += b (per column); H = silu(Z) + Hres (Hres may be NULL). */
int bt_gnn_bias_silu(float* Z, const float* b, float* H, const float* Hres, int32_t n, int32_t f, void* stream);
/* Y += b (per column). */
int bt_gnn_bias(float* Y, const float* b, int32_t n, int32_t f, void* stream);
/* dZ = dH * silu'(Z). */
int bt_gnn_silu_bwd(const float* dH, const float* Z, float* dZ, int64_t count, void* stream);
/* out_i = mean of H_j over the CSR neighbours of i (degree clamped to 1). */
int bt_gnn_mean_agg(const float* H, int64_t ldh, const int32_t* ptr, const int32_t* idx, float* out, int64_t ldo,
                    int32_t n, int32_t f, void* stream);
/* Adjoint of bt_gnn_mean_agg on a symmetric graph: out_j (+)= sum_{i in N(j)} G_i / deg_i. */
int bt_gnn_mean_agg_T(const float* G, int64_t ldg, const int32_t* ptr, const int32_t* idx, float* out,
                      int64_t ldo, int32_t n, int32_t f, int32_t accumulate, void* stream);
/* out_c = sum_r G[r][c], with per-block partials in ws (ws_len floats). */
int bt_gnn_colsum(const float* G, int64_t ldg, int32_t n, int32_t f, float* out, float* ws, int32_t ws_len,
                  void* stream);
/* bt_gnn_colsum without the workspace argument, allocating nothing: no
 * counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight training or serving loops, harnesses
 * that forbid hipMalloc). The per-block partial sums go to a file-scope device
 * array of BT_GNN_SCRATCH_FLOATS floats (1 MiB) that each GPU holds until the
 * process ends; its device address is looked up once per device, on the first
 * call with that device current (the current device must be the stream's).
 * The partials are every time written before they are read. Same result as
 * bt_gnn_colsum with ws_len = BT_GNN_SCRATCH_FLOATS; f > 256 returns
 * BT_EINVAL. Calls that overlap on one GPU (other streams or host threads)
 * share the array and race: serialise them. */
#define BT_GNN_SCRATCH_FLOATS (1 << 18)
int bt_gnn_colsum_noalloc(const float* G, int64_t ldg, int32_t n, int32_t f, float* out, void* stream);
/* dY = 2 (Y-T) mask inv_n; part[b] = partial sums of (Y-T)^2 mask inv_n per 256 elements. */
int bt_gnn_mse(const float* Y, const float* T, const float* mask, int64_t count, float inv_n, float* dY,
               float* part, int32_t part_len, void* stream);
/* One Adam update (step t >= 1). */
int bt_gnn_adam(float* p, const float* g, float* m, float* v, int64_t count, float lr, float b1, float b2,
                float eps, int32_t t, void* stream);

/* Wait for the stream; return BT_OK or BT_UNSTABLE (a NaN was produced; the
 * state is frozen at the step that produced it and *bad_step receives its
 * 0-based index counted from context creation, or -1). With streams > 1
 * (bt_tuning) the clusters of the other stream groups stop where their own
 * stream had got to within the same bt_step call. */
int bt_sync(bt_ctx* ctx, int64_t* bad_step);

/* Average device time in ms of one launch of a step kernel (0 membrane/cell,
 * 1 env field + diffusion (+ the exchange with exchange_form
 * BT_EXCHANGE_FUSED), 2 env exchange (nothing with BT_EXCHANGE_FUSED), 3 env
 * field only, 4 reaction network; 4 needs an attached network),
 * timed with HIP events over n_launches back-to-back launches on the context
 * stream. Leaves the state undefined: call after the timed work. */
int bt_time_kernel(bt_ctx* ctx, int32_t kernel, int32_t n_launches, double* ms_out);

/* Init residual (no counterpart in the reference, whose init runs a fixed
 * number of steps, parameters.py:525): r_out[q] = max over the cells of
 * cluster q and the ions of |c - c_prev| / |c|, c the current cc_cells and
 * c_prev a caller copy ([n_ions][n_cells]); cell_cluster[n_cells] maps cells
 * to clusters. r_out is device memory of n_clusters doubles. Asynchronous. */
int bt_residual(bt_ctx* ctx, const double* c_prev, const int32_t* cell_cluster, double* r_out);

/* Steps enqueued since bt_create. */
int64_t bt_steps_done(const bt_ctx* ctx);

/* Last error message of this context (or of the last failed bt_create). */
const char* bt_last_error(const bt_ctx* ctx);

/* ABI version (major*100+minor). */
int bt_version(void);

#ifdef __cplusplus
}
#endif

#endif /* BETSE_AMD_H */
