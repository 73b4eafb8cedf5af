/*
 * bigdft_hip.h -- C ABI of the MI355X local-Hamiltonian library.
 *
 * Drop-in replacement for the accelerator layer of BigDFT's local
 * Hamiltonian application (the GPU%OCLconv branch of
 * LocalHamiltonianApplication, bigdft/src/hpsiortho.f90:892-903, completed in
 * SynchronizeHamiltonianApplication, hpsiortho.f90:1313-1323).
 *
 * Conventions (kept from the reference):
 *   - keys are int32, 1-based: keyg(2,nseg) = first/last linear coarse index
 *     i3*(n1+1)*(n2+1) + i2*(n1+1) + i1 + 1, keyv(nseg) = 1-based offset into
 *     the coefficient vector (bigdft/src/locregs/initlocregs.f90:1633-1686);
 *     coarse and fine descriptors are passed separately, as allocate_data_OCL
 *     uploads them (bigdft/src/interface_ocl.f90:129-148);
 *   - one orbital = [psi_c(nvctr_c) ; psi_f(7, nvctr_f)] doubles
 *     (bigdft/src/wavelib/un-compress.f90:2187-2189);
 *   - potential: n1i*n2i*n3i doubles, x fastest, n1i = 2n+2 (periodic) or
 *     2n+31 (free) (bigdft/src/locregs/locregs.f90:871-884);
 *   - hpsi is ACCUMULATED (hpsi += H_loc psi), as compress_and_accumulate_*
 *     does on the CPU (un-compress.f90:2105-2238) and as the OpenCL path's
 *     result is added after finish (hpsiortho.f90:1316-1321);
 *   - ekin, epot are per orbital and unweighted: ekin = <psi|T|psi>,
 *     epot = sum V psir^2; the caller applies kwgt*occup as
 *     finish_hamiltonian_OCL does (interface_ocl.f90:589-607).
 * Every entry point returns 0 on success or a BDFT_ERR_* code; the message
 * of the last failure of a context is available from bdft_last_error.
 * Calls on one context are asynchronous with respect to the host until
 * bdft_sync (like the OpenCL queue, interface_ocl.f90:37).
 */
#ifndef BIGDFT_HIP_H
#define BIGDFT_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdft_ctx bdft_ctx;
typedef struct bdft_lr bdft_lr;
typedef struct bdft_plan bdft_plan;

/* Confining potential of one orbital: the fields of confpot_data
 * (bigdft/src/locregs/locreg_operations.f90:23-30).  The local potential
 * becomes V + prefac * r2^(potorder/2) (integer division), r2 the squared
 * distance of the ISF point hh*(i + ioffset) (1-based index i of the
 * region's fine grid) from rxyzConf, wherever the potential box covers the
 * region (apply_potential_lr_conf*, bigdft/src/orbitals/potential.f90:
 * 203-560).  potorder 0: no confinement.  damping is not used here. */
typedef struct bdft_confpot {
  int potorder;
  int ioffset[3];
  double prefac;
  double hh[3];
  double rxyzConf[3];
  double damping;
} bdft_confpot;

enum {
  BDFT_OK = 0,
  BDFT_ERR_ARG = 1,         /* null handle, bad size or geocode */
  BDFT_ERR_HIP = 2,         /* HIP runtime failure */
  BDFT_ERR_DESC = 3,        /* inconsistent keyg/keyv descriptors */
  BDFT_ERR_NOMEM = 4,       /* device allocation failed */
  BDFT_ERR_UNSUPPORTED = 5  /* box or mode outside the supported range */
};

/* arithmetic of the convolution passes */
enum {
  BDFT_PREC_FP64 = 0,    /* plain FP64 everywhere */
  BDFT_PREC_FP32 = 1,    /* FP32 convolution passes on the vector units, FP64
                            potential, energies and hpsi accumulation */
  BDFT_PREC_BF16X3 = 2,  /* split-BF16 MFMA, FP32 accumulation: filter and data
                            split hi+lo, products hi*hi + hi*lo + lo*hi */
  BDFT_PREC_BF16X4 = 3,  /* as BF16X3 plus the third filter piece times the data
                            hi part (removes the filter's fixed 2^-16 error) */
  BDFT_PREC_F16X3 = 4    /* x/y forward transforms and the kinetic on fp16
                            pieces (three products, 2^-22 representation):
                            data scaled by a power of two per box, each row
                            of the G tables and each x line of the G_x output
                            by its own power of two; z potential pass in
                            FP32; transposed transforms as BF16X4; FP32
                            accumulation */
};

/* Context on one device.  Replaces init_acceleration_OCL
 * (bigdft/src/interface_ocl.f90:20-41: ocl_create_context,
 * ocl_build_programs, ocl_create_command_queue_id). */
int bdft_ctx_create(int device, bdft_ctx **ctx);
/* Replaces free_gpu_OCL (interface_ocl.f90:196) for the context part.
 * Regions and captured plans still alive are released on the device and
 * detached: destroying them afterwards frees their host objects only, and
 * executing a detached plan is BDFT_ERR_ARG. */
int bdft_ctx_destroy(bdft_ctx *ctx);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL restores the
 * context's own stream.  Work queued later on the new stream waits for the
 * work already queued on the old one. */
int bdft_ctx_set_stream(bdft_ctx *ctx, void *hip_stream);
/* Message of the last failure on this context ("" if none).  Replaces the
 * print-and-exit of oclErrorCheck (bigdft/src/OpenCL/OpenCL_wrappers.h:89). */
const char *bdft_last_error(const bdft_ctx *ctx);

/* Localisation region: box, grid spacing and compression keys; keys are
 * validated and uploaded once.  Replaces allocate_data_OCL
 * (interface_ocl.f90:43-193).  geocode 'P' (periodic) or 'F' (free); n[d] is
 * the reference's n_d (coarse points n_d+1; the Hamiltonian
 * entry points take at most 256 per axis, the split precisions running
 * their FP32 vector passes on boxes beyond 64; the transfers at most
 * 4096); nfl/nfu
 * bound the fine region (ignored for 'P'); hgrid is the coarse spacing. */
int bdft_lr_create(bdft_ctx *ctx, char geocode, const int n[3],
                   const int nfl[3], const int nfu[3], const double hgrid[3],
                   int nseg_c, int nvctr_c, const int32_t *keyg_c,
                   const int32_t *keyv_c, int nseg_f, int nvctr_f,
                   const int32_t *keyg_f, const int32_t *keyv_f,
                   bdft_lr **lr);
/* Replaces free_gpu_OCL (interface_ocl.f90:196-269) for the descriptor part */
int bdft_lr_destroy(bdft_lr *lr);
/* n1i,n2i,n3i and nvctr_c + 7*nvctr_f of a region */
int bdft_lr_info(const bdft_lr *lr, int nfine[3], int64_t *ncoef);

/* Local Hamiltonian of norb orbitals sharing one region, device pointers:
 * psi, hpsi: norb consecutive orbitals; pot: nspin consecutive potentials
 * of n1i*n2i*n3i doubles (pot(n1i,n2i,n3i,nspin)); ekin, epot: norb doubles
 * each (device).  spinsgn: host array of norb doubles (orbs%spinsgn of the
 * batch) or NULL for all spin up; orbital i reads the first potential where
 * spinsgn[i] > 0 and the second one otherwise (nspin = 2; a spin-down
 * orbital with nspin = 1 is BDFT_ERR_ARG).  Replaces ocl_fulllocham_generic_k
 * (bigdft/src/OpenCL/Hamiltonian.c:105-172) called in the orbital loop of
 * local_hamiltonian_OCL (interface_ocl.f90:391-407 arguments, rhopot_up /
 * rhopot_down chosen per orbital at :487-491), all orbitals in one batch. */
int bdft_locham_d(bdft_ctx *ctx, const bdft_lr *lr, int norb, int nspin,
                  const double *pot, const double *spinsgn, const double *psi,
                  double *hpsi, double *ekin, double *epot, int precision);

/* Batched application over orbitals living in different regions (the ket
 * loop over localisation regions, hpsiortho.f90:934-958 with
 * modules/orbital_basis.f90:760-858).  Orbital i uses region lrs[i], its
 * coefficients at psi + psi_off[i] (hpsi likewise), and the potential box
 * pot + pot_off[i] of extents potdim[3i..3i+2] placed at ishift[3i..3i+2] in
 * the region's fine grid (apply_potential_lr_nobounds,
 * bigdft/src/orbitals/potential.f90:850-879).  Device pointers.  Host
 * arrays pot_off, potdim, ishift, psi_off may be NULL: potential offset 0,
 * extents n1i,n2i,n3i of the region, no shift, orbitals stored one after
 * the other. */
int bdft_locham_batch_d(bdft_ctx *ctx, int norb, const bdft_lr *const *lrs,
                        const double *pot, const int64_t *pot_off,
                        const int32_t *potdim, const int32_t *ishift,
                        const double *psi, const int64_t *psi_off,
                        double *hpsi, double *ekin, double *epot,
                        int precision);
/* Bytes of the fixed device array of bdft_locham_batch_fixed_d (one copy
 * per device the library runs on, held until the process ends). */
#define BDFT_FIXED_BYTES (128ull << 20)
/* bdft_locham_batch_d without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * kernels as bdft_locham_batch_d; the plan's descriptors, operator tables
 * and working grids live in a fixed device array of BDFT_FIXED_BYTES per
 * device, and the orbitals run in consecutive chunks that fit in it.  An
 * orbital whose working grids alone exceed it (8*(2*W + T2) bytes in FP64,
 * half in the other precisions, W = 8*N1*N2*N3 and T2 = n1i*n2i*2*N3) is
 * BDFT_ERR_UNSUPPORTED.  The call returns when the batch is complete.  All
 * calls on one device share the array: calls that overlap on one device
 * (several host threads, or contexts on different streams) race. */
int bdft_locham_batch_fixed_d(bdft_ctx *ctx, int norb, const bdft_lr *const *lrs,
                              const double *pot, const int64_t *pot_off,
                              const int32_t *potdim, const int32_t *ishift,
                              const double *psi, const int64_t *psi_off,
                              double *hpsi, double *ekin, double *epot,
                              int precision);
/* The same with a confining potential per orbital (conf[norb], host
 * array; the ket loop passes psi%confdata to psir_to_vpsi,
 * modules/orbital_basis.f90:830-831; the linear-scaling Hamiltonian of
 * the reference runs this on the CPU only).  epot includes the
 * confinement energy, as the reference's does. */
int bdft_locham_conf_batch_d(bdft_ctx *ctx, int norb, const bdft_lr *const *lrs,
                             const double *pot, const int64_t *pot_off,
                             const int32_t *potdim, const int32_t *ishift,
                             const double *psi, const int64_t *psi_off,
                             double *hpsi, double *ekin, double *epot,
                             const bdft_confpot *conf, int precision);

/* Density of norb orbitals sharing one region (device pointers psi, rho):
 * rho[s*n1i*n2i*n3i + p] += hfac[i] * psir_i(p)^2 for the orbitals i of
 * spin s, psir_i = daub_to_isf(psi_i) on the region's fine ISF grid
 * (bigdft/src/orbitals/daubisf.f90:13-80).  Replaces local_partial_density_OCL
 * (interface_ocl.f90:908-1040, ocl_locden_generic per orbital) and the CPU
 * local_partial_density / partial_density (bigdft/src/sumrho.f90:377-560,
 * npsir = 1, no slab scattering).  hfac (host, norb doubles) is
 * kwgt*occup/volume_element as the caller computes it (sumrho.f90:425);
 * spin (host, norb ints in [0,nspin), or NULL for all 0) picks the density
 * component as spinsgn does (sumrho.f90:515-519); nspin is 1 or 2.  rho is
 * accumulated, not initialised (the OpenCL layer pre-fills it with 1e-20,
 * interface_ocl.f90:957-960; here that is the caller's choice).  Orbitals are
 * summed in a fixed order, so results do not depend on scheduling.
 * precision: FP64 or FP32.  Returns after the density is complete. */
int bdft_density_d(bdft_ctx *ctx, const bdft_lr *lr, int norb,
                   const double *psi, const double *hfac, const int32_t *spin,
                   int nspin, double *rho, int precision);

/* bdft_density_d on host arrays psi (norb x ncoef) and rho (nspin x
 * n1i*n2i*n3i, accumulated): the argument form of local_partial_density_OCL
 * (interface_ocl.f90:908-921), which reads host psi and returns host rho_p.
 * Copies in, computes, copies back; returns when rho is updated. */
int bdft_density(bdft_ctx *ctx, const bdft_lr *lr, int norb, const double *psi,
                 const double *hfac, const int32_t *spin, int nspin, double *rho,
                 int precision);

/* Global -> local compressed transfer for norb orbitals (device pointers):
 * orbital i of the global region glr (at gpsi + goff[i], default
 * i*(nvctr_c+7*nvctr_f) of glr) into local region llrs[i] whose box starts at
 * ns[3i..3i+2] in glr's box (at lpsi + loff[i], default consecutive).
 * Replaces psi_to_locreg2 (locregs/locreg_operations.f90:2143-2318):
 * lpsi is set to the global values at the local points, zero where a local
 * point is not in glr.  Free-BC regions only, as shift_locreg_indexes
 * requires (:2338-2343).  Returns when done. */
int bdft_psi_to_locreg_d(bdft_ctx *ctx, const bdft_lr *glr, int norb,
                         const bdft_lr *const *llrs, const int32_t *ns,
                         const double *gpsi, const int64_t *goff, double *lpsi,
                         const int64_t *loff);

/* Local -> global: the values of each local orbital written at its points
 * of the global orbital (other global values untouched).  Replaces
 * Lpsi_to_global2 (locreg_operations.f90:1665-1837, nspin = 1); where the
 * reference stops because a local point is not in the global region
 * (:1827-1831) this returns BDFT_ERR_DESC.  Returns when done. */
int bdft_locreg_to_psi_d(bdft_ctx *ctx, const bdft_lr *glr, int norb,
                         const bdft_lr *const *llrs, const int32_t *ns,
                         const double *lpsi, const int64_t *loff, double *gpsi,
                         const int64_t *goff);

/* Preconditioner, the device part of preconditionall_OCL
 * (bigdft/src/interface_ocl.f90:622-880): ocl_preconditioner_generic_k
 * (bigdft/src/OpenCL/Preconditioner.c:245-358), real wavefunctions,
 * batched over orbitals.  For orbital i (region lrs[i], coefficients at
 * psi_off[i] of x and b; NULL psi_off: consecutive), ncong conjugate-
 * gradient steps on S (T + cprecr[i]) S y = b from the guess x, then
 * x = S y; T is the kinetic operator of the region (periodic, or free),
 * S the diagonal scaling 1/sqrt(sum_d {A2|B2}/(2 hgrid_d^2) + cprecr) of
 * each component type (OpenCL/Uncompress.c:377-470).  x and b are what
 * precondition_preconditioner (wfn_opt/precond.f90:673) leaves on the
 * host (the scaled guess and the scaled right-hand side); x, b device
 * pointers, cprecr host (cprecr_from_eval, precond.f90:472).  b is not
 * modified.  Returns when done. */
int bdft_precond_cg_d(bdft_ctx *ctx, int norb, const bdft_lr *const *lrs,
                      const int64_t *psi_off, const double *cprecr, int ncong,
                      double *x, const double *b);
/* The same for norb consecutive orbitals of one region, host arrays. */
int bdft_precondition(bdft_ctx *ctx, const bdft_lr *lr, int norb,
                      const double *cprecr, int ncong, double *x, const double *b);

/* Convolution bounds of a free-BC region: the 17 arrays of
 * convolutions_bounds (bigdft/src/locregs/bounds.f90:44-50) in one int32
 * buffer, in the order kb%ibyz_c ibxz_c ibxy_c ibyz_f ibxz_f ibxy_f,
 * gb%ibzxx_c ibxxyy_c ibyz_ff ibzxx_f ibxxyy_f, sb%ibzzx_c ibyyzz_c ibxy_ff
 * ibzzx_f ibyyzz_f, ibyyzz_r.  Array k starts at offset[k] (offset[17] is
 * the total) and is the Fortran array ib(2, lo1:hi1, lo2:hi2), column-major,
 * with dims[4k..4k+3] = lo1, hi1, lo2, hi2.  Pure host function. */
int bdft_bounds_layout(const int n[3], const int nfl[3], const int nfu[3],
                       int64_t offset[18], int32_t dims[68]);
/* Replaces locreg_bounds (bounds.f90:555-613; called through
 * ensure_locreg_bounds, bigdft/src/locregs/locregs.f90:664-676): the arrays
 * computed on the device from the region's keys into a device buffer of
 * offset[17] int32 (stream-ordered).  Values as the reference's, including
 * ibyz_c squared into 2x2 blocks (make_ib_c squares it in place).  Free BC
 * only, n <= 484 per axis (the reference's 1000 sentinel must stay above
 * every index). */
int bdft_locreg_bounds_d(bdft_ctx *ctx, const bdft_lr *lr, int32_t *bounds);
/* The same into a host buffer; returns when done. */
int bdft_locreg_bounds(bdft_ctx *ctx, const bdft_lr *lr, int32_t *bounds);

/* Plans: the same batch with fixed pointers, captured once into a hipGraph
 * and replayed per SCF iteration.  No counterpart in the reference (the
 * OpenCL path re-enqueues every kernel per orbital). */
int bdft_plan_create(bdft_ctx *ctx, int norb, const bdft_lr *const *lrs,
                     const double *pot, const int64_t *pot_off,
                     const int32_t *potdim, const int32_t *ishift,
                     const double *psi, const int64_t *psi_off, double *hpsi,
                     double *ekin, double *epot, int precision,
                     bdft_plan **plan);
int bdft_plan_execute(bdft_plan *plan);
int bdft_plan_destroy(bdft_plan *plan);

/* Host-array convenience with the argument meaning of local_hamiltonian_OCL
 * (interface_ocl.f90:391-407): pot(n1i,n2i,n3i,nspin), spinsgn(norb) (NULL:
 * all spin up) as bdft_locham_d; copies pot and psi in, hpsi (accumulated)
 * and the per-orbital energies out, then synchronises. */
int bdft_local_hamiltonian(bdft_ctx *ctx, const bdft_lr *lr, int norb, int nspin,
                           const double *pot, const double *spinsgn,
                           const double *psi, double *hpsi, double *ekin,
                           double *epot, int precision);

/* Compressed <-> mixed-layout transfers of norb orbitals of one region,
 * device pointers; psig holds norb consecutive (0:n1,2,0:n2,2,0:n3,2)
 * volumes.  bdft_uncompress_d zeroes psig and scatters psi (uncompress_per
 * without the synthesis, un-compress.f90:877-936; OpenCL uncompress_d,
 * bigdft/src/OpenCL/Uncompress.c:23-86); bdft_compress_acc_d does
 * hpsi += gather(psig) (compress_and_accumulate_mixed,
 * un-compress.f90:2174-2238).  Both are pure index work and bit-exact. */
int bdft_uncompress_d(bdft_ctx *ctx, const bdft_lr *lr, int norb,
                      const double *psi, double *psig);
int bdft_compress_acc_d(bdft_ctx *ctx, const bdft_lr *lr, int norb,
                        const double *psig, double *hpsi);

/* Stage timing with HIP events around every kernel of non-graph launches
 * (replaces the OpenCL event list, init_event_list/print_event_list,
 * bigdft/src/OpenCL/OpenCL_wrappers.h:309-311).  bdft_profile_read
 * synchronises, then returns the accumulated milliseconds of the first
 * nstage stages (names from bdft_stage_name, launch order) and the number of
 * profiled launches; reset != 0 clears the counters. */
int bdft_profile_enable(bdft_ctx *ctx, int on);
int bdft_profile_read(bdft_ctx *ctx, int nstage, double *ms, int64_t *nlaunch,
                      int reset);
const char *bdft_stage_name(int stage);
/* LDS bytes a work-group may use on the context's device, and whether the
 * plane-fused kernel sequence is enabled (environment BDFT_LINE_PASSES=1 at
 * context creation selects the line-pass sequence for every batch). */
int bdft_ctx_info(const bdft_ctx *ctx, int64_t *lds_per_block,
                  int *plane_enabled);

/* Launch options of the context.  No counterpart in the reference (its
 * OpenCL kernels have fixed work-group shapes); these are the choices the
 * autotuner (bigdft_hip/autotune.py) times per box shape and caches.
 *   BDFT_OPT_LINE_PASSES  0: plane-fused sequences where they fit (default),
 *                         1: line passes (FP64/FP32) and per-pass
 *                            matrix-core kernels for every batch
 *   BDFT_OPT_ZGROUP       matrix-core z-line bundles of 32 lines per
 *                         work-group: 0 automatic (3 unless the z kernels
 *                         would then take more than half a CU's LDS), 2, 3
 *                         (3 falls back to 2 where it does not fit)
 *   BDFT_OPT_ZPOT         z potential pass (G_z, V, G_z^T) of the
 *                         matrix-core precisions: 0 automatic (the FP32
 *                         vector-unit line kernel where every box extent is
 *                         at most 64 coarse points, else the matrix cores),
 *                         1 matrix cores (split precision; F16X3 keeps the
 *                         FP32 pass), 2 FP32 vector units where they apply
 *   BDFT_OPT_BWD          backward plane pass (G_y^T G_x^T + K, compress)
 *                         of the matrix-core precisions: 0 automatic (the
 *                         FP32 vector-unit plane kernel for free-BC
 *                         batches, the matrix cores otherwise), 1 matrix
 *                         cores, 2 FP32 vector units where the plane kernel
 *                         applies (its planes fit in LDS)
 *   BDFT_OPT_CHUNK_MB     Hamiltonian batches and plans: 0 one launch per
 *                         stage over the whole batch; n > 0 the orbitals in
 *                         consecutive chunks whose working grids take at
 *                         most n MiB, each chunk's stages run before the
 *                         next chunk's on one workspace of that size, so
 *                         the grids a stage writes are read back from the
 *                         last-level cache (the density, preconditioner and
 *                         allocation-free calls are not chunked)
 *   BDFT_OPT_TILE_ORDER   0: the plane and (un)compress passes take their
 *                         tiles in dispatch order (default); 1: in one
 *                         contiguous range per XCD, as the z-line and line
 *                         passes always do (work-groups are dealt to the 8
 *                         XCDs in turn, each with its own L2)
 *   BDFT_OPT_PERSIST      0: the z potential pass launches one work-group
 *                         per tile (default); 1: a persistent launch of
 *                         resident work-groups over a work queue of the
 *                         tiles, boxes in order of decreasing size (each
 *                         work-group ends at the end of its stretch)
 * A change drops the context's cached plans; plans already created by
 * bdft_plan_create keep the options they were made with. */
#define BDFT_OPT_LINE_PASSES 1
#define BDFT_OPT_ZGROUP 2
#define BDFT_OPT_ZPOT 3
#define BDFT_OPT_BWD 4
#define BDFT_OPT_CHUNK_MB 5
#define BDFT_OPT_TILE_ORDER 6
#define BDFT_OPT_PERSIST 7
int bdft_ctx_set_option(bdft_ctx *ctx, int option, int value);
int bdft_ctx_get_option(const bdft_ctx *ctx, int option, int *value);

/* Wait for all work of the context.  Replaces ocl_finish in
 * finish_hamiltonian_OCL (interface_ocl.f90:599). */
int bdft_sync(bdft_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
