/*
 * dsph_abi.h -- C ABI of the MI355X-native DualSPHysics SPH step.
 *
 * This is the drop-in boundary of the GPU SPH step. Every entry point replaces
 * one host-callable operator of the reference's CUDA layer (namespaces cusph,
 * cusphs, cudiv; reference v5.4.355, paths relative to src/source/), keeping the
 * argument meaning, the particle-array layout and the call order of the
 * reference host driver (JSphGpuSingle.cpp / JSphGpu.cpp / JCellDivGpu*.cpp).
 *
 * Differences from the reference interface, all deliberate:
 *  - constants live in a per-context struct (dsph_cte) instead of the global
 *    __constant__ CTE (JSphGpu_ker.cu:39), so several independent cases can
 *    share one device (batched ensembles);
 *  - every call returns a status code (DSPH_OK, ...) instead of relying on
 *    Check_CudaErroor + JException (RunExceptionGpuDef.h:19-35); the message of
 *    the last failure is in dsph_last_error();
 *  - every call is asynchronous on the context's stream, except the ones whose
 *    name ends in _sync or that return values to the host (documented per call);
 *  - no entry point allocates device memory except dsph_malloc and
 *    dsph_ctx_reserve; scratch (sort buffers, reduction partials) belongs to
 *    the context and is sized by dsph_ctx_reserve.
 *
 * Layout contract (identical to the reference, JSphGpu.h:103-209):
 *  SoA arrays indexed [0,Npb) boundary, [Npb,Np) fluid, sorted by cell after
 *  every divide. Positions are FP64 (double2 posxy + double posz); poscell is
 *  float4{x,y,z, bits(PSCEL code)} relative to a cell of size poscellsize;
 *  velrho is float4{vx,vy,vz,rho}; code is u16 with the bit fields of
 *  DualSphDef.h:150-241; dcell is u32 with the DCEL coding of JDsDcellDef.h.
 *  Symmetric tensors (tsymatrix3f) are 6 floats {xx,xy,xz,yy,yz,zz}.
 *
 * All pointers named d_* are device pointers; h_* are host pointers.
 */
#ifndef DSPH_ABI_H
#define DSPH_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSPH_ABI_VERSION 1

/* Status codes. */
#define DSPH_OK        0
#define DSPH_E_ARG     1   /* invalid argument (null pointer, size, mode) */
#define DSPH_E_LAUNCH  2   /* kernel launch or runtime error */
#define DSPH_E_NOMEM   3   /* device allocation failed or scratch too small */
#define DSPH_E_STATE   4   /* call not valid in the current context state */
#define DSPH_E_NUMERIC 5   /* NaN/Inf dt, boundary particle excluded, ... */

/* Option enums: values equal the reference's (DualSphDef.h). */
#define DSPH_KERNEL_CUBIC     1   /* TpKernel KERNEL_Cubic (not built) */
#define DSPH_KERNEL_WENDLAND  2   /* TpKernel KERNEL_Wendland */
#define DSPH_VISCO_ARTIFICIAL 1   /* TpVisco */
#define DSPH_VISCO_LAMINARSPS 2
#define DSPH_VISCO_LAMINAR    3
#define DSPH_DDT_NONE         0   /* TpDensity */
#define DSPH_DDT_DDT          1   /* Molteni & Colagrossi */
#define DSPH_DDT_DDT2         2   /* Fourtakas (inner particles) */
#define DSPH_DDT_DDT2FULL     3   /* Fourtakas (full) */
#define DSPH_BC_DBC           1   /* TpBoundary */
#define DSPH_BC_MDBC          2
#define DSPH_SLIP_VEL0        1   /* TpSlipMode: mDBC, boundary velocity 0 */
#define DSPH_SLIP_NOSLIP      2   /* mDBC2 no-slip (MDBC2_Std) */
#define DSPH_SLIP_FREESLIP    3   /* mDBC2 free slip */
#define DSPH_SHIFT_NONE       0   /* TpShifting */
#define DSPH_SHIFT_NOBOUND    1   /* no shifting near any boundary */
#define DSPH_SHIFT_NOFIXED    2   /* no shifting near fixed boundaries */
#define DSPH_SHIFT_FULL       3
#define DSPH_BMODE_DBC        0   /* boundmode of a boundary particle */
#define DSPH_BMODE_MDBC2      1
#define DSPH_BMODE_MDBC2OFF   2   /* ghost node not submerged: mass off */

/* Particles per interaction, simulation or ensemble: the interaction kernels
 * address the float4 particle arrays with 32-bit byte offsets (16 bytes per
 * particle), so np must stay below 2^28; larger counts are refused with
 * DSPH_E_ARG. */
#define DSPH_NP_LIMIT 268435456u

typedef struct dsph_ctx dsph_ctx;

/*
 * Constants of the interaction, mirror of StCteInteraction
 * (JSphGpu_cte.h:27-62), filled as JSphGpu::GetConstantData (JSphGpu.cpp:564-588).
 */
typedef struct dsph_cte {
  float massb, massf;            /* reference masses [kg] */
  float kernelh;                 /* smoothing length h [m] */
  float kernelsize2;             /* (2h)^2 */
  float poscellsize;             /* size of PosCell cells (usually 2h) */
  float awen, bwenh;             /* Wendland constants (FunSphKernel.h:198-207) */
  float cs0;                     /* speed of sound at rho0 */
  float eta2;                    /* (0.1h)^2 */
  float ddtkh;                   /* DDTValue*KernelSize */
  float ddtgz;                   /* rho0*|g.z|/B */
  float scell;                   /* cell size of the neighbour search */
  float kernelsize;              /* 2h */
  float dp;                      /* inter-particle spacing */
  float cteb;                    /* B of Tait's EOS [Pa] */
  float gamma;                   /* polytropic constant */
  float rhopzero;                /* rho0 */
  float ovrhopzero;              /* 1/rho0 */
  float movlimit;                /* 0.9*Scell, max displacement per step */
  uint32_t tboundary;            /* DSPH_BC_* */
  uint32_t cellcode;             /* DomCellCode (JSph.cpp:1936) */
  double domposminx, domposminy, domposminz;
  double maprealposminx, maprealposminy, maprealposminz;
  double maprealsizex, maprealsizey, maprealsizez;
  /* Periodic boundaries (JSph.cpp:810-826, 2183-2194): bits 1 X, 2 Y, 4 Z;
     the increment applied to a particle leaving through the upper face of
     each periodic axis (PeriXinc, PeriYinc, PeriZinc). */
  uint32_t periactive;
  uint32_t simulate2d;           /* 2-D simulation (Simulate2D, JSph.cpp:646): the
                                    mDBC systems drop the y row and column */
  double xperinc[3], yperinc[3], zperinc[3];
} dsph_cte;

/*
 * Cell-division data for the neighbour search, mirror of StDivDataGpu
 * (JCellDivDataGpu.h:27-79) for single-GPU (axis Z).
 */
typedef struct dsph_divdata {
  int32_t scelldiv;              /* 1 (Full cells) or 2 (Half cells) */
  int32_t ncx, ncy, ncz, ncw;    /* cells of the current domain, ncw=ncx*ncy */
  uint32_t cellfluid;            /* first fluid cell = ncw*ncz+1 */
  int32_t cellzerox, cellzeroy, cellzeroz; /* CellDomainMin */
  const int32_t* d_beginendcell; /* int2[SizeBeginEndCell] {begin,end} */
  float scell;
  uint32_t domcellcode;
  double domposminx, domposminy, domposminz;
  float kernelsize2;
  float poscellsize;
} dsph_divdata;

/*
 * Parameters of one force interaction, mirror of the subset of StInterParmsg
 * (JSphGpu_ker.h:55-203) used by the in-scope options (no floatings, no
 * shifting, no periodic, no inlet/outlet).
 */
typedef struct dsph_inter_parms {
  int32_t tkernel;               /* DSPH_KERNEL_WENDLAND */
  int32_t tvisco;                /* DSPH_VISCO_* */
  int32_t tdensity;              /* DSPH_DDT_* */
  float viscob, viscof;          /* Visco*ViscoBoundFactor, Visco */
  uint32_t np, npb, npbok;       /* Np, Npb, NpbOk */
  dsph_divdata divdata;
  const uint32_t* d_dcell;
  const float*    d_poscell;     /* float4[np] */
  const float*    d_velrho;      /* float4[np] */
  const uint16_t* d_code;
  const float*    d_spstaurho2;  /* sym6[np] (VISCO_LAMINARSPS) else NULL */
  float* d_viscdt;               /* float[np]   (in/out, max-accumulated) */
  float* d_ar;                   /* float[np]   (in/out, summed) */
  float* d_ace;                  /* float3[np]  (in/out, summed) */
  float* d_sps2strain;           /* sym6[np]    (in/out, VISCO_LAMINARSPS) */
  float* d_velgrad;              /* float[9*np] full velocity gradient (learned
                                    closure features) or NULL */
  const uint8_t* d_boundmode;    /* mDBC2: DSPH_BMODE_* of [0,npb), else NULL */
  const float*   d_tangenvel;    /* mDBC2: float3[npb] tangential velocity */
  int32_t shiftmode;             /* DSPH_SHIFT_* (with d_shiftposfs) */
  float* d_shiftposfs;           /* shifting: float4[np] (in/out; fluid part
                                    cleared before the interaction), else NULL */
  const float* d_ftomassp;       /* floating bodies: mass of the particles of each
                                    body, indexed by the code's type value
                                    (FtoMasspg of StInterParmsg), else NULL */
  const float* d_boundnor;       /* mDBC2 NoPenetration: float3[np] boundary normals
                                    (BoundNor of StInterParmsg), else NULL */
  float* d_nopenshift;           /* mDBC2 NoPenetration (MDBC2_NoPen): float4[np] out,
                                    written for every fluid particle: velocity
                                    correction per axis, w = 10 when any applies
                                    (NoPenShift of StInterParmsg), else NULL */
  const float* d_motionvel;      /* mDBC2 NoPenetration with moving boundaries:
                                    float3[np] motion velocity of the boundary
                                    particles (MotionVel of StInterParmsg), the
                                    no-penetration test uses the velocity relative
                                    to it; NULL = fixed boundaries */
} dsph_inter_parms;

/* ---------------------------------------------------------------- context */

/* Creates a context on HIP device `device` with its own stream. */
int dsph_ctx_create(int device, dsph_ctx** out);
int dsph_ctx_destroy(dsph_ctx* ctx);
/* Sizes context scratch for up to np_max particles and nctt_max cells
 * (sort double-buffers, digit histograms, reduction partials). Allocates. */
int dsph_ctx_reserve(dsph_ctx* ctx, uint32_t np_max, uint32_t nctt_max);
/* Replaces the context stream (hipStream_t as void*); NULL = own stream. */
int dsph_ctx_set_stream(dsph_ctx* ctx, void* stream);
void* dsph_ctx_get_stream(dsph_ctx* ctx);
int dsph_synchronize(dsph_ctx* ctx);
const char* dsph_last_error(dsph_ctx* ctx);
/* Device memory (mirror of the JArraysGpu pool's allocations). */
int dsph_malloc(dsph_ctx* ctx, size_t bytes, void** d_out);
int dsph_free(dsph_ctx* ctx, void* d_ptr);
int dsph_memset_async(dsph_ctx* ctx, void* d_ptr, int value, size_t bytes);
int dsph_memcpy_h2d(dsph_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int dsph_memcpy_d2h(dsph_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
int dsph_memcpy_d2d_async(dsph_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* Per-kernel timing with HIP events recorded on the context stream around
 * every kernel launch of the entry points (mirror of the reference's stage
 * timers, JDsTimersGpu.h / JTimerCuda.h). Off by default; collect resolves
 * the pending events (synchronizes) into per-kernel totals; get iterates the
 * totals (name, launches, total ms); reset clears them. */
int dsph_ctx_set_timing(dsph_ctx* ctx, int enabled);
int dsph_ctx_timing_collect(dsph_ctx* ctx);
int dsph_ctx_timing_get(dsph_ctx* ctx, uint32_t idx, char* name, size_t name_len,
                        uint64_t* count, double* total_ms);
int dsph_ctx_timing_reset(dsph_ctx* ctx);

/* Context options (no counterpart in the reference: implementation choices
 * with identical results up to FP32 rounding). */
#define DSPH_OPT_FORCES_KERNEL 1  /* fluid interaction kernel: 0 = cell rows
                                     staged in LDS per wave with per-lane pair
                                     queues (default), 1 = one lane per
                                     particle over its cell rows */
#define DSPH_OPT_BOUND_KERNEL  2  /* boundary interaction kernel (with
                                     DSPH_OPT_FORCES_KERNEL 0): 0 = auto (the
                                     default: times both forms on interactions
                                     1 and 3 of every 4096 and keeps the faster;
                                     one host synchronisation per trial), 1 =
                                     staged rows (tile), 2 = one lane per
                                     particle. Identical bits either way. */
int dsph_ctx_set_option(dsph_ctx* ctx, int key, int value);

/* Replaces cusph::CteInteractionUp (JSphGpu_ker.h:284). Copies into ctx. */
int dsph_set_constants(dsph_ctx* ctx, const dsph_cte* cte);

/* ----------------------------------------------------- neighbour list (cudiv) */

/* Replaces cudiv::LimitsCell (JCellDivGpu_ker.h:51): min/max cell (x,y,z) of
 * the non-excluded particles [pini,pini+np). Synchronous (host results).
 * With no valid particle, celmin > celmax (as the reference). */
int dsph_cell_limits_sync(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                          uint32_t cellcode, const uint32_t* d_dcell,
                          const uint16_t* d_code,
                          uint32_t h_celmin[3], uint32_t h_celmax[3]);

/* Replaces cudiv::PreSortFull (JCellDivGpuSingle_ker.h:28): cell key of every
 * particle (boxes of JCellDivGpuSingle.cpp:109-128) and sortpart[p]=p. */
int dsph_presort_full(dsph_ctx* ctx, uint32_t np, uint32_t cellcode,
                      const uint32_t* d_dcell, const uint16_t* d_code,
                      const uint32_t cellmin[3], const uint32_t ncells[3],
                      uint32_t* d_cellpart, uint32_t* d_sortpart);

/* Replaces cudiv::PreSortFluid (JCellDivGpuSingle_ker.h:31): fluid particles
 * [pini,pini+npf) only. */
int dsph_presort_fluid(dsph_ctx* ctx, uint32_t npf, uint32_t pini,
                       uint32_t cellcode, const uint32_t* d_dcell,
                       const uint16_t* d_code, const uint32_t cellmin[3],
                       const uint32_t ncells[3], uint32_t* d_cellpart,
                       uint32_t* d_sortpart);

/* ---- Periodic boundaries (cusph, JSphGpu_ker.cu:2106-2400) ---- */
/* Replaces cusph::PeriodicIgnore (JSphGpu_ker.cu:2113-2132): marks the
 * periodic copies of [0,n) CODE_OUTIGNORE. */
int dsph_periodic_ignore(dsph_ctx* ctx, uint32_t n, uint16_t* d_code);
/* Replaces cusph::PeriodicMakeList (JSphGpu_ker.cu:2142-2213): list of the
 * particles of [pini,pini+n) (normal or periodic codes) whose position plus
 * perinc (entry p) or minus perinc (entry p|0x80000000) falls inside
 * [mapposmin,mapposmax). The list is in particle order, the plus entry of a
 * particle before its minus entry (the order of the CPU build,
 * JSphCpuSingle.cpp:210-240). d_work holds ceil(n/256)+1 words; at most nmax
 * entries are written and the full count is returned in *h_count (host
 * memory; the call synchronises). */
int dsph_periodic_make_list(dsph_ctx* ctx, uint32_t n, uint32_t pini, uint32_t nmax,
                            const double mapposmin[3], const double mapposmax[3],
                            const double perinc[3], const double* d_posxy, const double* d_posz,
                            const uint16_t* d_code, uint32_t* d_work, uint32_t* d_listp,
                            uint32_t* h_count);
/* Replaces cusph::PeriodicDuplicateSymplectic / PeriodicDuplicateVerlet /
 * PeriodicDuplicateNormals (JSphGpu_ker.cu:2229-2400): creates particles
 * [pini,pini+n) from the list: position +-perinc, cell from the constants'
 * DomPosMin clamped to domcells-1, the same id, a periodic code, velrho and,
 * when not null, the Symplectic pre state, the Verlet velrhom1, the SPS tau
 * and the mDBC normals. */
int dsph_periodic_duplicate(dsph_ctx* ctx, uint32_t n, uint32_t pini, const uint32_t domcells[3],
                            const double perinc[3], const uint32_t* d_listp, uint32_t* d_idp,
                            uint16_t* d_code, uint32_t* d_dcell, double* d_posxy, double* d_posz,
                            float* d_velrho, float* d_spstau, double* d_posxypre,
                            double* d_poszpre, float* d_velrhopre, float* d_velrhom1,
                            float* d_boundnor);

/* Replaces cudiv::Sort (JCellDivGpu_ker.h:40): in-place sort of (keys,values)
 * by key, keys < key_limit. Always stable (the reference's -stable mode):
 * LSD radix sort, ceil(log2(key_limit)/8) passes. Uses context scratch. */
int dsph_sort_by_cell(dsph_ctx* ctx, uint32_t* d_keys, uint32_t* d_values,
                      uint32_t n, uint32_t key_limit);

/* Replaces cudiv::CalcBeginEndCell (JCellDivGpu_ker.h:53): clears
 * begcell[0..sizebegcell) (full) or [cellfluid..sizebegcell) and writes
 * {first,last+1} of every non-empty cell from the sorted keys. */
int dsph_begin_end_cell(dsph_ctx* ctx, int full, uint32_t np, uint32_t npb,
                        uint32_t sizebegcell, uint32_t cellfluid,
                        const uint32_t* d_cellpart, int32_t* d_begcell);

/* Replaces the CellBeginEnd readback (JCellDivGpuSingle.cpp:218-226): copies
 * ncells int2 entries from cell `cellini` to the host. Synchronous. */
int dsph_cell_begin_end_sync(dsph_ctx* ctx, const int32_t* d_begcell,
                             uint32_t cellini, uint32_t ncells,
                             int32_t* h_out);

/* Replaces cudiv::SortDataParticles (basic overload, JCellDivGpu_ker.h:56;
 * called from JCellDivGpu::SortBasicArrays, JCellDivGpu.cpp:329-337):
 * gathers [pini,np) of idp, code, dcell, posxy, posz, velrho by sortpart. */
int dsph_sort_data_basic(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                         const uint32_t* d_sortpart,
                         const uint32_t* d_idp, const uint16_t* d_code,
                         const uint32_t* d_dcell, const double* d_posxy,
                         const double* d_posz, const float* d_velrho,
                         uint32_t* d_idp2, uint16_t* d_code2, uint32_t* d_dcell2,
                         double* d_posxy2, double* d_posz2, float* d_velrho2);
/* Same gather fused with UpdatePosCell of the gathered positions
 * (JSphGpuSingle.cpp:396-406 followed by :476-477). poscell covers [0,np):
 * entries below pini are recomputed from the (unmoved) posxy/posz. */
int dsph_sort_data_basic_poscell(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                         const uint32_t* d_sortpart,
                         const uint32_t* d_idp, const uint16_t* d_code,
                         const uint32_t* d_dcell, const double* d_posxy,
                         const double* d_posz, const float* d_velrho,
                         uint32_t* d_idp2, uint16_t* d_code2, uint32_t* d_dcell2,
                         double* d_posxy2, double* d_posz2, float* d_velrho2,
                         const double posmin[3], float poscellsize,
                         float* d_poscell);
/* Other SortDataParticles overloads (JCellDivGpu_ker.h:58-76). */
int dsph_sort_data_f4(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                      const uint32_t* d_sortpart, const float* d_a, float* d_a2);
int dsph_sort_data_f3(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                      const uint32_t* d_sortpart, const float* d_a, float* d_a2);
int dsph_sort_data_sym6(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                        const uint32_t* d_sortpart, const float* d_a, float* d_a2);
int dsph_sort_data_d2_d_f4(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                           const uint32_t* d_sortpart,
                           const double* d_a, const double* d_b, const float* d_c,
                           double* d_a2, double* d_b2, float* d_c2);

/* Replaces cusphs::UpdatePosCell (JSphGpuSimple_ker.h:34). */
int dsph_update_poscell(dsph_ctx* ctx, uint32_t np, const double posmin[3],
                        float poscellsize, const double* d_posxy,
                        const double* d_posz, float* d_poscell);

/* ------------------------------------------------------ interaction (cusph) */

/* Replaces cusph::Interaction_Forces (JSphGpu_ker.h:293): fluid-fluid and
 * fluid-boundary pass over fluid particles [npb,np) and boundary-fluid pass
 * over [0,npbok). Outputs are accumulated (ace,ar summed; viscdt max), as in
 * KerInteractionForcesFluid / KerInteractionForcesBound. */
int dsph_interaction_forces(dsph_ctx* ctx, const dsph_inter_parms* t);

/* Particles per call of dsph_interaction_forces_noalloc: its scratch (24
 * bytes per particle, 24 MiB) is a static device array that every GPU the
 * library runs on holds until the process ends. */
#define DSPH_NOALLOC_NP_MAX 1048576u

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dsph_interaction_forces with the same arguments and the
 * same kernels and results, without dsph_ctx_reserve and without any
 * allocation; the per-particle preparation (pressure, 1/rho, map-frame
 * positions) lives in the library's static device scratch. np above
 * DSPH_NOALLOC_NP_MAX is refused with DSPH_E_ARG. All calls on one GPU share
 * that GPU's scratch: calls that overlap on one GPU (other streams, other
 * contexts, other threads) race. Kernel timing (dsph_ctx_set_timing) creates
 * HIP events and must stay off for a call that may not allocate. */
int dsph_interaction_forces_noalloc(dsph_ctx* ctx, const dsph_inter_parms* t);

/* Replaces cusph::ComputeSpsTau (JSphGpu_ker.h:304). */
int dsph_sps_tau(dsph_ctx* ctx, uint32_t np, uint32_t npb, float smag,
                 float blin, const float* d_velrho, const float* d_sps2strain,
                 float* d_tau);

/* ---------------------------------------------------------------- learned closure
 * A per-particle MLP (11 features -> 64 -> 64 -> 1, SiLU, softplus output)
 * predicting the eddy viscosity of the SPS stress from the velocity gradient
 * (9, row-major d v_a / d x_b), density and dp. No counterpart in the
 * reference: it replaces the Smagorinsky viscosity smag*sqrt(prr) of
 * cusph::ComputeSpsTau (JSphGpu_ker.h:304) when enabled. Products run on the
 * matrix cores in BF16 with FP32 accumulation; parameters are FP32 masters
 * with BF16 copies. Buffers are owned by the caller. */
typedef struct dsph_mlp {
  float* d_params;               /* FP32 parameters [dsph_mlp_param_count()] */
  uint16_t* d_wbf16;             /* BF16 weight copies [dsph_mlp_bf16_count()] */
  float* d_adam_m;               /* Adam moments [param_count] (training only) */
  float* d_adam_v;
  float feat_mean[11], feat_scale[11];  /* x' = (x - mean) * scale */
  float out_scale;               /* nu = out_scale * softplus(output) [m^2/s] */
} dsph_mlp;
/* Parameter block layout: W1[64][16] (input columns 11..15 unused), b1[64],
 * W2[64][64], b2[64], w3[64], b3. */
int dsph_mlp_param_count(void);
int dsph_mlp_bf16_count(void);
/* Refreshes the BF16 copies from the FP32 parameters. */
int dsph_mlp_sync(dsph_ctx* ctx, const dsph_mlp* m);
/* d_out[i] = nu(features[i]), features FP32 [n][11]. */
int dsph_mlp_forward(dsph_ctx* ctx, const dsph_mlp* m, uint32_t n,
                     const float* d_features, float* d_out);
/* Gradient of L = mean_i (softplus(o_i) - t_i/out_scale)^2 over the n
 * samples into d_grad [param_count] and L into d_loss (may be NULL).
 * d_scratch holds per-block partials: at least 21256 bytes, more gives more
 * parallelism (up to 1024 blocks). Deterministic (fixed reduction order). */
int dsph_mlp_grad(dsph_ctx* ctx, const dsph_mlp* m, uint32_t n,
                  const float* d_features, const float* d_targets,
                  float* d_scratch, size_t scratch_bytes, float* d_grad,
                  float* d_loss);
/* One Adam step (step = 1, 2, ...) on the FP32 parameters, then the BF16
 * copies. */
int dsph_mlp_adam(dsph_ctx* ctx, const dsph_mlp* m, const float* d_grad,
                  float lr, float beta1, float beta2, float eps, uint32_t step);
/* SPS tau of the fluid particles [npb,np) with the learned eddy viscosity
 * (features from d_velgrad [9*np], rho, dp); the blin term as
 * cusph::ComputeSpsTau. d_nu [np] receives nu_t (may be NULL). */
int dsph_sps_tau_learned(dsph_ctx* ctx, const dsph_mlp* m, uint32_t np,
                         uint32_t npb, float blin, float dp,
                         const float* d_velrho, const float* d_velgrad,
                         const float* d_sps2strain, float* d_tau, float* d_nu);
/* Training data from a fine-resolution state: Shepard/Wendland filter with
 * smoothing length hf at nprobe probe points (FP64 [nprobe][3]) over the
 * fluid particles; 23 FP64 moments per probe (see learned.hip). */
int dsph_filter_probes(dsph_ctx* ctx, uint32_t nprobe, const double* d_probes,
                       double hf, const dsph_divdata* dvd, float massf,
                       const double* d_posxy, const double* d_posz,
                       const float* d_velrho, const uint16_t* d_code,
                       double* d_out);

/* Replaces cusph::Interaction_MdbcCorrection (JSphGpu_mdbc_iker.h:32),
 * SlipMode SLIP_Vel0 (mDBC with vel=0): density of boundary particles
 * [0,n) extrapolated from their ghost node; writes velrho[p].w. */
int dsph_mdbc_correction(dsph_ctx* ctx, uint32_t n, const dsph_divdata* dvd,
                         const double mapposmin[3], const double* d_posxy,
                         const double* d_posz, const float* d_poscell,
                         const uint16_t* d_code, const float* d_boundnor,
                         float* d_velrho);

/* Replaces cusph::Interaction_Mdbc2Correction (JSphGpu_mdbc_iker.h:44),
 * SlipMode SLIP_NoSlip with MDBC2_KEEPVEL (DualSphDef.h:471): for boundary
 * particles [0,n) with a normal, density at the ghost node (4x4 first-order
 * correction when well conditioned, else Shepard), pressure cloned to the
 * particle, velrho[p].w written; d_boundmode[p] = DSPH_BMODE_MDBC2 or
 * DSPH_BMODE_MDBC2OFF (dry ghost node) and d_tangenvel[p] = tangential part
 * of 2 v_motion - v_ghost. The caller clears d_boundmode[0,n) to
 * DSPH_BMODE_DBC first (JSphGpuSingle.cpp:700-701). d_motionvel and
 * d_motionace (float3[n]) may be NULL for fixed boundaries. */
int dsph_mdbc2_correction(dsph_ctx* ctx, uint32_t n, const float gravity[3],
                          const dsph_divdata* dvd, const double mapposmin[3],
                          const double* d_posxy, const double* d_posz,
                          const float* d_poscell, const uint16_t* d_code,
                          const float* d_boundnor, const float* d_motionvel,
                          const float* d_motionace, float* d_velrho,
                          uint8_t* d_boundmode, float* d_tangenvel);
/* The same with the slip mode of Interaction_Mdbc2Correction
 * (JSphGpu_mdbc_iker.cu:289-487, 517-560): DSPH_SLIP_NOSLIP (as above) or
 * DSPH_SLIP_FREESLIP, for which the tangential velocity is the tangential
 * part of the ghost node's Shepard velocity. */
int dsph_mdbc2_correction_slip(dsph_ctx* ctx, int32_t slipmode, uint32_t n,
                               const float gravity[3], const dsph_divdata* dvd,
                               const double mapposmin[3], const double* d_posxy,
                               const double* d_posz, const float* d_poscell,
                               const uint16_t* d_code, const float* d_boundnor,
                               const float* d_motionvel, const float* d_motionace,
                               float* d_velrho, uint8_t* d_boundmode, float* d_tangenvel);

/* Replaces the no-penetration blocks of cusphs::ComputeStepSymplecticCor and
 * cusphs::ComputeStepVerlet (JSphGpuSimple_ker.cu:177-191, 447-461; mDBC2
 * with NoPenetration, MDBC2_NoPen): for fluid particles [npb,np) whose
 * d_nopenshift (float4[np], from dsph_interaction_forces) has w > 5, each
 * component c with a non-zero correction gets velocity d_velbase.c +
 * nopenshift.c and displacement that velocity times dt. Run after the step
 * kernel, on its outputs (d_velrho, d_movxy, d_movz); d_velbase = the
 * velocity the step starts from (velrhopre for the Symplectic corrector,
 * the current velocity for Verlet). dt = d_dt[dt_slot] when d_dt is not NULL
 * (device-side dt), else dt. */
int dsph_mdbc2_nopen_step(dsph_ctx* ctx, uint32_t np, uint32_t npb,
                          const float* d_nopenshift, const float* d_velbase, double dt,
                          const double* d_dt, uint32_t dt_slot, float* d_velrho,
                          double* d_movxy, double* d_movz);

/* Replaces cusph::CopyMotionVelAce (JSphGpu_mdbc_iker.cu:587-613; called at
 * the end of JSphGpu::RunMotion, JSphGpu.cpp:1219-1223) for mDBC2 with moving
 * boundaries: the moving particles among the boundary particles [0,npb)
 * (idp in [idbegin, idbegin+count)) get d_motionace = float((v -
 * d_motionvel)/dt) in FP64 and d_motionvel = v (float3[npb] each, zero at
 * the start, reordered with the particles). Run after the step's motion,
 * whether the motion was active or not. dt = d_dt[dt_slot] when d_dt is not
 * NULL, else dt. */
int dsph_mdbc2_motion_vel_ace(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin, uint32_t count,
                              double dt, const double* d_dt, uint32_t dt_slot,
                              const uint32_t* d_idp, const float* d_velrho, float* d_motionvel,
                              float* d_motionace);

/* Replaces cusph::Interaction_GaugeVel and the mass sums of
 * Interaction_GaugeSwl (JDsGauge_ker.cu:50-245) for n points at once
 * (double3[n]): d_vel (float3[n]) = sum over fluid within 2h of W*V*v
 * (no Shepard normalisation, as the reference), d_mass (float[n]) =
 * sum W*V*massf, FP64 sums; points outside [mapmin, mapmax) give 0. Either
 * output may be NULL. The SWL level follows from the masses along the
 * gauge line (JGaugeSwl, JDsGaugeItem.cpp:758-790). */
int dsph_gauge_points(dsph_ctx* ctx, uint32_t n, const double* d_points,
                      const dsph_divdata* dvd, const double mapmin[3],
                      const double mapmax[3], const double* d_posxy,
                      const double* d_posz, const float* d_velrho,
                      const uint16_t* d_code, float* d_vel, float* d_mass);
/* Replaces cusph::Interaction_GaugeMaxz (JDsGauge_ker.cu:252-312) for n
 * gauges: point0 (double3[n]) and {height, distlimit} (float2[n]); d_zmax =
 * max(highest fluid z within distlimit of the vertical line, zfloor). */
int dsph_gauge_maxz(dsph_ctx* ctx, uint32_t n, const double* d_points,
                    const float* d_height_dist, const dsph_divdata* dvd,
                    double zfloor, const double* d_posxy, const double* d_posz,
                    const uint16_t* d_code, float* d_zmax);

/* Replaces cusph::Interaction_GaugeForce (JDsGauge_ker.cu:497-570): for
 * the boundary particles [0,npb) with id in [idbegin, idbegin+count),
 * d_ace[id-idbegin] (float3[count], cleared first) = pressure acceleration
 * from the fluid, sum of -(P1+P2)/(rho1 rho2) massf gradW; the gauge's force
 * is massbound times the sum over the object (JDsGaugeItem.cpp:1895-1900). */
int dsph_gauge_force(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin, uint32_t count,
                     const dsph_divdata* dvd, const double* d_posxy,
                     const double* d_posz, const uint16_t* d_code,
                     const uint32_t* d_idp, const float* d_velrho, float* d_ace);

/* Replaces cushift::RunShifting (JSphShifting_ker.cu:123-165; host
 * JSphShifting::RunGpu JSphShifting.cpp:463-470) for fluid [pini,pini+n):
 * umagn = dt*shiftcoef*kernelh*|v| (times (w-tfs)/(3-tfs) above the
 * free-surface threshold tfs, 0 below; tfs = 0 disables it; 2-tfs when the
 * context's constants have simulate2d set), shift =
 * shiftposfs.xyz*umagn clamped to +-0.1 dp; a cancelled particle (x =
 * FLT_MAX) gets 0. dt = d_dt[dt_slot]*dt_scale when d_dt is not NULL. */
int dsph_run_shifting(dsph_ctx* ctx, uint32_t n, uint32_t pini, double dt,
                      const double* d_dt, uint32_t dt_slot, double dt_scale,
                      double shiftcoef, double kernelh, float shifttfs, double dp,
                      const float* d_velrho, float* d_shiftposfs);

/* --------------------------------------------------- reductions and dt */

/* Replaces cusph::ComputeVelMod + ReduMaxFloat (JSphGpu_ker.h:281,290):
 * max over [pini,pini+n) of |v|^2 written to *d_out (device float). */
int dsph_velmax2(dsph_ctx* ctx, uint32_t n, uint32_t pini,
                 const float* d_velrho, float* d_out);
/* Replaces cusph::ComputeAceMod + ReduMaxFloat (JSphGpu_ker.h:287): max of
 * |ace|^2 over [pini,pini+n). */
int dsph_acemax2(dsph_ctx* ctx, uint32_t n, uint32_t pini,
                 const float* d_ace, float* d_out);
/* Replaces cusph::ComputeAceMod with codes + ReduMaxFloat (JSphGpu_ker.cu:260-275):
 * max of |ace|^2 over the normal particles of [pini,pini+n) (periodic copies
 * left out); d_work holds n floats. */
int dsph_acemax2_normal(dsph_ctx* ctx, uint32_t n, uint32_t pini, const float* d_ace,
                        const uint16_t* d_code, float* d_work, float* d_out);
/* Replaces cusph::ReduMaxFloat (JSphGpu_ker.h:281): max of data[ini..ini+n)
 * (0 when n==0) to *d_out. */
int dsph_redu_max_f32(dsph_ctx* ctx, uint32_t n, uint32_t ini,
                      const float* d_data, float* d_out);
/* Synchronous variant returning the value to the host, as ReduMaxFloat. */
int dsph_redu_max_f32_sync(dsph_ctx* ctx, uint32_t n, uint32_t ini,
                           const float* d_data, float* h_out);

/* Replaces JSphGpu::DtVariable (JSphGpu.cpp:1103-1147) on the device so the
 * step needs no host round trip. Reads d_red[0]=max|v|^2, d_red[1]=ViscDtMax,
 * d_red[2]=max|ace|^2 and writes d_dt[slot] (FP64):
 *   dt1=sqrt(h/AceMax), dt2=h/(max(cs0,10*VelMax)+h*ViscDtMax),
 *   dt=max(cfl*min(dt1,dt2), dtmin); NaN/Inf sets d_flags bit 0. */
int dsph_dt_variable(dsph_ctx* ctx, const float* d_red, double kernelh,
                     double cs0, double cfl, double dtmin, double* d_dt,
                     uint32_t slot, uint32_t* d_flags);

/* End of a Symplectic step on the device (JSphGpuSingle.cpp:819), with
 * d_dt = {dt_p, dt_c, dtpre, time, dtlast}: time += dtpre; dtlast = dtpre;
 * dtpre = min(dt_p, dt_c). */
int dsph_dt_advance(dsph_ctx* ctx, double* d_dt);

/* ------------------------------------------------- integration (cusphs) */

/* d_shiftposfs (shifting, else NULL): the shifting displacement
 * (dsph_run_shifting) is added to the fluid displacement
 * (JSphGpuSimple_ker.cu:164-168, 303-308, 460-465).
 * d_boundmode (mDBC2, else NULL): boundary particles with mode >=
 * DSPH_BMODE_MDBC2 keep velrhopre (Verlet: velrho2) instead of the DBC
 * density update (JSphGpuSimple_ker.cu:142,276,406-412). */
/* Replaces cusphs::ComputeStepSymplecticPre (JSphGpuSimple_ker.h:51) for
 * DBC/mDBC without floatings/shifting/inout. dtm = dt/2. When d_dt is not
 * NULL, dtm is read on the device as 0.5*d_dt[dt_slot] instead. */
int dsph_step_symplectic_pre(dsph_ctx* ctx, uint32_t np, uint32_t npb,
                             const float* d_velrhopre, const uint8_t* d_boundmode,
                             const float* d_ar,
                             const float* d_ace, const float* d_shiftposfs, double dtm,
                             const double* d_dt, uint32_t dt_slot,
                             float rhopzero, float rhopoutmin, float rhopoutmax,
                             const float gravity[3], uint16_t* d_code,
                             double* d_movxy, double* d_movz, float* d_velrho);
/* Replaces cusphs::ComputeStepSymplecticCor (JSphGpuSimple_ker.h:57).
 * dtm = dt/2; device dt as above (dt = d_dt[dt_slot]). */
int dsph_step_symplectic_cor(dsph_ctx* ctx, uint32_t np, uint32_t npb,
                             const float* d_velrhopre, const uint8_t* d_boundmode,
                             const float* d_ar,
                             const float* d_ace, const float* d_shiftposfs, double dtm,
                             double dt,
                             const double* d_dt, uint32_t dt_slot,
                             float rhopzero, float rhopoutmin, float rhopoutmax,
                             const float gravity[3], uint16_t* d_code,
                             double* d_movxy, double* d_movz, float* d_velrho);
/* Replaces cusphs::ComputeStepVerlet (JSphGpuSimple_ker.h:46). */
int dsph_step_verlet(dsph_ctx* ctx, uint32_t np, uint32_t npb,
                     const float* d_velrho1, const float* d_velrho2,
                     const uint8_t* d_boundmode,
                     const float* d_ar, const float* d_ace, const float* d_shiftposfs,
                     double dt, double dt2,
                     float rhopzero, float rhopoutmin, float rhopoutmax,
                     const float gravity[3], uint16_t* d_code,
                     double* d_movxy, double* d_movz, float* d_velrhonew);
/* Replaces cusph::ComputeStepPos2 (JSphGpu_ker.h:315): x=xpre+mov in FP64
 * for fluid [npb,np), OUTMOV/OUTPOS/NaN checks and new dcell. */
int dsph_step_pos2(dsph_ctx* ctx, uint32_t np, uint32_t npb,
                   const double* d_posxypre, const double* d_poszpre,
                   const double* d_movxy, const double* d_movz,
                   double* d_posxy, double* d_posz, uint32_t* d_dcell,
                   uint16_t* d_code);
/* Replaces cusph::ComputeStepPos (JSphGpu_ker.h:312): in-place update. */
int dsph_step_pos(dsph_ctx* ctx, uint32_t np, uint32_t npb,
                  const double* d_movxy, const double* d_movz,
                  double* d_posxy, double* d_posz, uint32_t* d_dcell,
                  uint16_t* d_code);

/* Imposed motion of one moving boundary object, active for
 * start <= t < finish; over a step [t, t+dt]:
 *  type 1, rectilinear sinusoid (JMotionMovRectSinu, JMotionMov.h:237-254;
 *    law JMotionObj.cpp:558-575): each axis with a non-zero amplitude moves
 *    by ampl*(sin(phase+2*pi*freq*(t+dt)) - sin(phase+2*pi*freq*t)), velocity
 *    = displacement/dt (JMotionList.cpp:43-50) -> MoveLinBound;
 *  type 2, rotation sinusoid (JMotionMovRotSinu, law JMotionObj.cpp:577-587):
 *    rotation about the axis axisp1->axisp2 by
 *    angampl*(sin(angphase+2*pi*angfreq*(t+dt)) - sin(angphase+2*pi*angfreq*t))
 *    degrees -> MoveMatBound;
 *  type 3, rotation at constant angular velocity (JMotionMovRot, law
 *    JMotionObj.cpp:518-524): by angvel*dt degrees -> MoveMatBound.
 * Rotations follow JMatrix4::MatrixRot (JMatrix4.h:465-497): a positive
 * angle turns clockwise seen from axisp2 towards axisp1's side, i.e. by
 * -angle about the direction axisp2-axisp1 (right-hand rule). */
typedef struct dsph_motion {
  int32_t type;                  /* 0 none, 1 rect. sinusoid, 2 rot. sinusoid, 3 rotation */
  uint32_t idbegin, count;       /* particle ids of the object */
  double freq[3], ampl[3], phase[3];  /* type 1: 1/s, m, rad */
  double start, finish;          /* s */
  double axisp1[3], axisp2[3];   /* types 2, 3: rotation axis points (m) */
  double angfreq, angampl, angphase;  /* type 2: 1/s, degrees, rad */
  double angvel;                 /* type 3: degrees/s */
} dsph_motion;

/* ------------------------------------------------------ floating bodies */

/* Replaces cusph::CalcRidp (JSphGpu_ker.h, KerCalcRidp JSphGpu_ker.cu:1582-1628)
 * for the floating ids: d_ridp[id - idini] = p for the particles p of
 * [pini, pini+n) with ids in [idini, idfin) (the others are not written). */
int dsph_ft_calc_ridp(dsph_ctx* ctx, uint32_t n, uint32_t pini, uint32_t idini,
                      uint32_t idfin, const uint32_t* d_idp, uint32_t* d_ridp);
/* Replaces cusph::FtPartsSumAce (JSphGpu_ker.cu:1938-2012): for each body cf
 * (d_ftbody = {first index into d_ridp, count} per body, d_center double3 per
 * body), d_acelinang[6cf..6cf+5] = the sum of ace and of (pos - centre) x ace
 * over its particles, in the reference kernel's order (256 strided partial
 * sums, then summed in order). */
int dsph_ft_parts_sum_ace(dsph_ctx* ctx, uint32_t ftcount, const uint32_t* d_ftbody,
                          const double* d_center, const uint32_t* d_ridp,
                          const double* d_posxy, const double* d_posz, const float* d_ace,
                          float* d_acelinang);
/* Replaces cusph::FtPartsUpdate (JSphGpu_ker.cu:2027-2092) without periodic
 * axes or mDBC normals: the count particles d_ridp[fpini..] of one body move
 * by dt times their velocity (KerUpdatePos: position, cell, out-of-domain
 * code) and take the rigid velocity fvel + fomega x (pos - fcenter). */
int dsph_ft_parts_update(dsph_ctx* ctx, double dt, uint32_t count, uint32_t fpini,
                         const float fvel[3], const float fomega[3], const double fcenter[3],
                         const uint32_t* d_ridp, double* d_posxy, double* d_posz,
                         float* d_velrho, uint32_t* d_dcell, uint16_t* d_code);

/* Replaces cusph::MoveLinBound (JSphGpu_ker.h:326): displaces the moving
 * boundary particles with id in [idbegin, idbegin+count) by mvpos (FP64,
 * KerUpdatePos rules) and sets their velocity to mvvel. The particles are
 * found by scanning the boundary block [0,npb) by id instead of the
 * reference's id->position map (RidpMot). */
int dsph_move_lin_bound(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin,
                        uint32_t count, const double mvpos[3],
                        const float mvvel[3], const uint32_t* d_idp,
                        double* d_posxy, double* d_posz, uint32_t* d_dcell,
                        float* d_velrho, uint16_t* d_code);
/* Replaces cusph::MoveMatBound (JSphGpu_ker.h:328; KerMoveMatBound
 * JSphGpu_ker.cu:1705-1740): the moving boundary particles with id in
 * [idbegin, idbegin+count) go to x' = M x (FP64, M row-major 3x4: rotation
 * and translation), displaced with the KerUpdatePos rules, velocity
 * (x'-x)/dt; their mDBC normals (d_boundnor, float3, may be NULL) become
 * M(x+n) - M(x). */
int dsph_move_mat_bound(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin,
                        uint32_t count, const double mat[12], double dt,
                        const uint32_t* d_idp, double* d_posxy, double* d_posz,
                        uint32_t* d_dcell, float* d_velrho, uint16_t* d_code,
                        float* d_boundnor);
/* The same with the displacement (type 1) or matrix (types 2, 3) of motion
 * m computed on the device from the device time slots (dt = d_dt[2],
 * t = t0 + d_dt[3], the layout of dsph_dt_variable): no host round trip per
 * step. No counterpart in the reference (its motion is computed on the host
 * every step). d_boundnor as in dsph_move_mat_bound. */
int dsph_move_motion_device(dsph_ctx* ctx, uint32_t npb, const dsph_motion* m,
                            double t0, const double* d_dt,
                            const uint32_t* d_idp, double* d_posxy,
                            double* d_posz, uint32_t* d_dcell, float* d_velrho,
                            uint16_t* d_code, float* d_boundnor);
/* The 3x4 FP64 matrix of a rotation motion (types 2, 3) over [t, t+dt]
 * (host computation, the reference's JMotionObj::ProcesTime path). */
int dsph_motion_matrix(const dsph_motion* m, double t, double dt, double mat[12]);

/* ------------------------------------------------------------ measurement */

/* Replaces cusph::InteractionNg (JDsPips_ker.cu:54-124) + reductions: counts
 * real pairs (rr2<=(2h)^2 and rr2>=1e-18) and checked candidates, as
 * {real fluid, real bound, checked fluid, checked bound}. Synchronous. */
int dsph_pips_count_sync(dsph_ctx* ctx, uint32_t nb, uint32_t pinitb,
                         uint32_t nf, uint32_t pinitf, const dsph_divdata* dvd,
                         const uint32_t* d_dcell, const float* d_poscell,
                         uint64_t h_res[4]);
/* Same count with the fluid real pairs also split: h_res = {real fluid,
 * real bound, checked fluid, checked bound, real fluid-boundary subset}
 * (the algorithmic-work model prices fluid-fluid and fluid-boundary pairs
 * differently). Synchronous. */
int dsph_pips_count_split_sync(dsph_ctx* ctx, uint32_t nb, uint32_t pinitb,
                               uint32_t nf, uint32_t pinitf, const dsph_divdata* dvd,
                               const uint32_t* d_dcell, const float* d_poscell,
                               uint64_t h_res[5]);

/* ----------------------------------------------------- whole-step driver */

/*
 * C++ host driver (mirror of JSphGpuSingle) exposed for bindings. A
 * simulation owns its particle arrays and calls only the entry points above.
 */
typedef struct dsph_sim dsph_sim;

/* A floating body of the case: mirror of the <floating> block of the case
 * particles (JCaseParts.cpp:274-297) as JSph::LoadCaseConfig keeps it in
 * StFloatingData (JSph.cpp:1123-1167). Its particles are the ids
 * [begin, begin+count) within [casenfixed+casenmoving, casenbound). */
typedef struct dsph_floating {
  uint32_t begin, count;         /* first id and number of particles */
  double massbody, masspart;     /* body mass and mass of each of its particles [kg] */
  double center[3];              /* centre of mass [m] */
  double inertia[9];             /* inertia tensor about the centre, row-major [kg m^2] */
  double velini[3], omegaini[3]; /* initial linear [m/s] and angular [rad/s] velocity */
} dsph_floating;

/* Case description: mirror of the <execution><constants> and <parameters>
 * blocks (JCaseCtes.cpp:208-222, JSph.cpp:663-854) plus the map limits. */
typedef struct dsph_case {
  double dp, kernelh, cteb, gamma, rhopzero, cs0;
  double massfluid, massbound;
  double gravity[3];
  double cflnumber;
  double dtini, dtmin;           /* 0 = derive from h/cs0 (JSph.cpp:1535-1538) */
  int32_t tstep;                 /* 1 Verlet, 2 Symplectic */
  int32_t verletsteps;
  int32_t tkernel, tvisco, tdensity, tboundary;
  double visco, viscoboundfactor, ddtvalue;
  double rhopoutmin, rhopoutmax;
  int32_t cellmode;              /* 1 Full, 2 Half */
  int32_t celldomfixed;          /* 0: adapt to particles (reference default) */
  double maprealposmin[3], maprealposmax[3];
  uint32_t casenp, casenbound, casenfixed, casenmoving, casenfloat, casenfluid;
  dsph_motion motion;            /* the moving object (ids casenfixed.. +casenmoving) */
  int32_t slipmode;              /* mDBC SlipMode (JSph.cpp:706-716): DSPH_SLIP_VEL0,
                                    DSPH_SLIP_NOSLIP or DSPH_SLIP_FREESLIP (the last two:
                                    fixed boundaries only) */
  int32_t shiftmode;             /* Shifting (JSph.cpp:744-760): DSPH_SHIFT_* */
  double shiftcoef, shifttfs;    /* ShiftCoef (-2 default), ShiftTFS (0: off) */
  int32_t periactive;            /* periodic axes, bits 1 X, 2 Y, 4 Z (JSph.cpp:810-826) */
  double perinc[3][3];           /* XPeriodicIncY/Z, YPeriodicIncX/Z, ZPeriodicIncX/Y: the
                                    increments [axis][component]; each axis' own
                                    component is set to -MapRealSize (JSph.cpp:2186-2188) */
  uint32_t ftcount;              /* floating bodies (casenfloat particles in all) */
  int32_t nopenetration;         /* mDBC2 NoPenetration (JSph.cpp:715-716): 1 = on */
  const dsph_floating* floatings; /* ftcount entries, in id order; read at creation */
  int32_t simulate2d;            /* <data2d> (JCaseCtes.cpp:209-211): particles in the
                                    plane y = simulate2dposy, 2-D Wendland constants
                                    (FunSphKernel.h:198-210) */
  int32_t pad_2d;
  double simulate2dposy;
} dsph_case;

/* State of a floating body (StFloatingData, DualSphDef.h): centre, rotation
 * angles, linear and angular velocity and acceleration. */
typedef struct dsph_floating_state {
  double center[3];
  float angles[3], fvel[3], fomega[3], facelin[3], faceang[3];
} dsph_floating_state;

/* Per-step statistics, mirror of the Run.out/Run.csv key numbers. */
typedef struct dsph_stats {
  uint32_t np, npb, npbok, nct;
  uint32_t nstep;
  uint32_t npf_out_total;        /* fluid particles excluded so far */
  double timestep;               /* simulated time */
  double dt_last;                /* dt of the last step */
  double velmax, acemax, viscdtmax;
  uint32_t ndiv, ndivfull;
  double symplecticdtpre;        /* dt of the next Symplectic step (saved in PARTs) */
} dsph_stats;

/* Creates a simulation from particles with the boundary block first: sorted
 * by idp (JPartsLoad4 order of a new case) or in the order a PART saved them
 * (a resumed run). h_pos is double3[np], h_velrho float4[np], h_boundnor
 * float3[casenbound] indexed by id, or NULL. */
int dsph_sim_create(int device, const dsph_case* cs, uint32_t np,
                    const uint32_t* h_idp, const double* h_pos,
                    const float* h_velrho, const float* h_boundnor,
                    dsph_sim** out);
int dsph_sim_destroy(dsph_sim* sim);
/* Time state of a run resumed from a PART file (JSph::InitRun, PartBegin
 * branch, JSph.cpp:2214-2219; JPartsLoad4.cpp:222): simulated time, the dt
 * of the next Symplectic step (<= 0 keeps DtIni) and the step count. */
int dsph_sim_set_time(dsph_sim* sim, double timestep, double symplecticdtpre,
                      uint32_t nstep);
const char* dsph_sim_last_error(dsph_sim* sim);
/* Runs nsteps full time steps (ComputeStep + RunCellDivide). With
 * sync_dt=0 the step keeps dt on the device (no host round trip). */
int dsph_sim_run(dsph_sim* sim, uint32_t nsteps, int sync_dt);
int dsph_sim_stats(dsph_sim* sim, dsph_stats* out);
/* Downloads current particles (sorted order): idp, pos (double3), velrho,
 * code. Any pointer may be NULL. Returns the count in *np. */
int dsph_sim_download(dsph_sim* sim, uint32_t* np, uint32_t* h_idp,
                      double* h_pos, float* h_velrho, uint16_t* h_code);
/* Downloads ace (float3), ar of the last interaction (sorted order). */
int dsph_sim_download_forces(dsph_sim* sim, float* h_ace, float* h_ar);
/* Counts pairs of the current state (PIPS rule), as
 * dsph_pips_count_split_sync: {real fluid, real bound, checked fluid,
 * checked bound, real fluid-boundary}. */
int dsph_sim_pips(dsph_sim* sim, uint64_t h_res[5]);
/* State of floating body cf after the last step (StFloatingData). */
int dsph_sim_floating(dsph_sim* sim, uint32_t cf, dsph_floating_state* out);
/* Enables the learned SPS closure (Laminar+SPS only; m = NULL disables):
 * the interaction also computes the velocity gradient and the SPS tau uses
 * the network's eddy viscosity (dsph_sps_tau_learned). The network buffers
 * stay owned by the caller and must outlive the simulation's use of them. */
int dsph_sim_set_closure(dsph_sim* sim, const dsph_mlp* m);
/* Eddy viscosity [np] of the last learned-closure evaluation and, when
 * h_velgrad is not NULL, the velocity gradient [np][9] (dv_a/dx_b, row-major)
 * it was evaluated on (sorted order, as dsph_sim_download). */
int dsph_sim_closure_nu(dsph_sim* sim, float* h_nu, float* h_velgrad);
/* Training data of the closure: dsph_filter_probes over the current state at
 * nprobe host probe points; h_out receives 23 FP64 moments per probe. */
int dsph_sim_filter_probes(dsph_sim* sim, uint32_t nprobe, const double* h_probes,
                           double hf, double* h_out);
/* Evaluates PreInteraction_Forces + Interaction_Forces on the current state
 * without advancing it (forces then readable with dsph_sim_download_forces). */
int dsph_sim_interact(dsph_sim* sim);
/* Gauges on the current state (after the last step's divide): h_points
 * double3[n] -> h_vel float3[n] and/or h_mass float[n] (dsph_gauge_points);
 * maxz gauges: h_points double3[n], h_height_dist float2[n] -> h_zmax
 * (dsph_gauge_maxz, floor = the map's minimum z). Synchronous. */
int dsph_sim_gauge_points(dsph_sim* sim, uint32_t n, const double* h_points,
                          float* h_vel, float* h_mass);
int dsph_sim_gauge_maxz(dsph_sim* sim, uint32_t n, const double* h_points,
                        const float* h_height_dist, float* h_zmax);
/* Force gauge of the boundary object with ids [idbegin, idbegin+count):
 * h_ace float3[count] per particle (dsph_gauge_force). Synchronous. */
int dsph_sim_gauge_force(dsph_sim* sim, uint32_t idbegin, uint32_t count, float* h_ace);
/* Access to the context of a simulation (streams, events). */
dsph_ctx* dsph_sim_ctx(dsph_sim* sim);

/*
 * Batched ensembles (no counterpart in the reference, which runs one case
 * per process): many independent cases advanced together on one GPU, each
 * stage of the step one launch for all cases, 24 steps per HIP graph launch
 * (3 for a remainder), the fluid interaction optionally one persistent
 * launch per bucket of case sizes. Every
 * case runs JSphGpuSingle's Symplectic step with DBC (artificial or laminar
 * viscosity, DDT2Full or no density diffusion, the same options for all
 * cases), a cell domain fixed to its map and its own adaptive dt, and stops
 * at its end time. A case's results equal a dsph_sim run of that case with
 * celldomfixed=1 and device-side dt, bit for bit.
 */
typedef struct dsph_ens dsph_ens;

typedef struct dsph_ens_state {
  uint32_t np, npbok;            /* current particles (excluded ones not counted) */
  uint32_t active;               /* 0 once time >= tend or after an error */
  uint32_t nstep;                /* steps run by this case */
  uint32_t flags;                /* 1 dt NaN/inf, 2 dt clamped to DtMin, 4 boundary excluded */
  uint32_t p0, np0;              /* range of the case in the concatenated arrays */
  uint32_t pad;
  double time, dt_last, dt_next;
} dsph_ens_state;

/* Creates an ensemble of ncases cases. Particles of all cases are
 * concatenated in case order (each case sorted by idp, boundary first, as for
 * dsph_sim_create): np[k] particles of case k, idp local to the case;
 * tend[k] = end time of case k. */
int dsph_ens_create(int device, uint32_t ncases, const dsph_case* cases,
                    const uint32_t* np, const uint32_t* h_idp, const double* h_pos,
                    const float* h_velrho, const double* tend, dsph_ens** out);
int dsph_ens_destroy(dsph_ens* e);
const char* dsph_ens_last_error(dsph_ens* e);
/* Runs steps in graph launches of 24 steps (of 3 for the last < 24 steps
 * of max_steps; eager launches with the graph disabled) until every case has
 * stopped or max_steps steps ran (rounded down to a multiple of 3); the stop
 * flags are read back once at least 3 * check_every steps ran since the last
 * reading. h_steps = steps run. */
int dsph_ens_run(dsph_ens* e, uint32_t max_steps, uint32_t check_every, uint32_t* h_steps);
int dsph_ens_set_graph(dsph_ens* e, int enable);
int dsph_ens_case_state(dsph_ens* e, uint32_t icase, dsph_ens_state* out);
/* Current particles of case icase in sorted order (as dsph_sim_download). */
int dsph_ens_download(dsph_ens* e, uint32_t icase, uint32_t* h_count, uint32_t* h_idp,
                      double* h_pos, float* h_velrho, uint16_t* h_code);
/* Autotuner: on each case-size bucket (fluid particles < 16384, < 131072,
 * larger), over the current state, reps runs each, it times the fluid-kernel
 * variants (LDS tile kernel with 8, 12 or 16 mask tiles per wave,
 * one-lane-per-particle kernel) on a grid of one workgroup per wave of 64
 * particles, then the best tile variant as a persistent launch (chunks of 1,
 * 2 or 4 waves of particles per work-counter fetch, grids of 10, 20 or 40
 * waves per CU), and the boundary kernel's shape (one lane per particle in
 * blocks of 64, 128, 256, or the tile form); it
 * keeps the fastest shapes. h_variant[3] (-1 for an empty bucket) and
 * h_ms[3*4] (best time of each variant on the plain grid) may be NULL; the
 * chosen shapes are read with dsph_ens_get_shape. */
int dsph_ens_autotune(dsph_ens* e, uint32_t reps, int32_t* h_variant, float* h_ms);
/* Plain grid with the given variant for a bucket (persistent form off). */
int dsph_ens_set_variant(dsph_ens* e, uint32_t bucket, int32_t variant);

/* Launch shape of a bucket's fluid interaction and of the boundary kernel. */
typedef struct dsph_ens_shape {
  int32_t variant;          /* 0..2 tile ring 8/12/16, 3 one lane per particle */
  uint32_t persistent;      /* 1: persistent launch (tile variants only) */
  uint32_t chunk;           /* persistent: waves of 64 particles per counter fetch (1..64) */
  uint32_t waves_per_cu;    /* persistent: grid = waves_per_cu x compute units (1..64) */
  uint32_t bound_block;     /* boundary kernel block size: 64, 128 or 256 (all buckets) */
  uint32_t bound_tile;      /* 1: boundary kernel in the tile form (bound_block 64) */
  float ms_plain;           /* autotuner: best plain-grid time of the bucket [ms] (0 untuned) */
  float ms_persistent;      /* autotuner: best persistent time [ms] (0 untuned) */
  float ms_bound;           /* autotuner: boundary kernel time with bound_block [ms] */
} dsph_ens_shape;
int dsph_ens_get_shape(dsph_ens* e, uint32_t bucket, dsph_ens_shape* out);
/* Sets a bucket's shape (ms fields ignored); bound_block applies to the
 * boundary kernel of all buckets. */
int dsph_ens_set_shape(dsph_ens* e, uint32_t bucket, const dsph_ens_shape* shape);
/* {cases, particles, sort keys, cases per bucket [3], steps per graph launch, steps run} */
int dsph_ens_info(dsph_ens* e, uint32_t* h_info);
int dsph_ens_set_timing(dsph_ens* e, int enable);
dsph_ctx* dsph_ens_context(dsph_ens* e);

#ifdef __cplusplus
}
#endif
#endif /* DSPH_ABI_H */
