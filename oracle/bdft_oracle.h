/*
 * CPU oracle for BigDFT's local-Hamiltonian path on Daubechies-16 wavelets.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's FP64 Fortran routines.  Only the test-suite, the smoke entry
 * and the CPU-baseline leg of the benchmark may load it, as the checker.
 * The product (the HIP library) never calls it.
 *
 * Array conventions follow the reference: Fortran column-major storage,
 * 1-based keys (keyg, keyv) exactly as produced by segkeys
 * (bigdft/src/locregs/initlocregs.f90:1633-1686).
 */
#ifndef BDFT_ORACLE_H
#define BDFT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* confpot_data (bigdft/src/locregs/locreg_operations.f90:23-30) */
typedef struct {
  int potorder;       /* 0: no confinement */
  int ioffset[3];     /* ISF offset of the box in the global region */
  double prefac;
  double hh[3];       /* ISF grid spacings (hgrid/2) */
  double rxyzConf[3]; /* centre, global coordinates */
  double damping;     /* not used by the local Hamiltonian */
} oref_confpot;

/* ---- filter tables (copied values; see bdft_oracle.c for the citation) ---- */
void oref_filters(double ch[16], double cg[16], double magic[16],
                  double kin_a[15], double kin_c[29], double kin_e[15],
                  double fil2[62]);

/* ---- periodic 1D rotating kernels (convolut_optim_per.f90) ---- */
void oref_ana_rot_per(int n, int ndat, const double *x, double *y);
void oref_syn_rot_per(int n, int ndat, const double *x, double *y);
void oref_convrot_n_per(int n1, int ndat, const double *x, double *y);
void oref_convrot_t_per(int n1, int ndat, const double *x, double *y);
void oref_convolut_kinetic_per_T(int n1, int n2, int n3, const double hgrid[3],
                                 const double *x, double *y, double kstrten[3]);

/* ---- periodic 3D drivers (convolut_common_per.f90) ---- */
void oref_synthese_per_self(int n1, int n2, int n3, double *x, double *y);
void oref_analyse_per_self(int n1, int n2, int n3, double *y, double *x);
void oref_convolut_magic_n_per(int n1, int n2, int n3, const double *x,
                               double *y, double *ww);
void oref_convolut_magic_t_per_self(int n1, int n2, int n3, double *x,
                                    double *y);

/* ---- compression (wavelib/un-compress.f90) ---- */
void oref_uncompress_per(int n1, int n2, int n3, int nseg_c, int nvctr_c,
                         const int32_t *keyg_c, const int32_t *keyv_c,
                         int nseg_f, int nvctr_f, const int32_t *keyg_f,
                         const int32_t *keyv_f, const double *psi_c,
                         const double *psi_f, double *psifscf, double *psig);
void oref_uncompress_psig(int n1, int n2, int n3, int nseg_c,
                          const int32_t *keyg_c, const int32_t *keyv_c,
                          int nseg_f, const int32_t *keyg_f,
                          const int32_t *keyv_f, const double *psi_c,
                          const double *psi_f, double *psig);
void oref_compress_and_accumulate_mixed(int n1, int n2, int n3, int nseg_c,
                                        const int32_t *keyg_c,
                                        const int32_t *keyv_c, int nseg_f,
                                        const int32_t *keyg_f,
                                        const int32_t *keyv_f,
                                        const double *psig, double *psi_c,
                                        double *psi_f);
void oref_uncompress_forstandard(int n1, int n2, int n3, const int nfl[3],
                                 const int nfu[3], int nseg_c,
                                 const int32_t *keyg_c, const int32_t *keyv_c,
                                 int nseg_f, const int32_t *keyg_f,
                                 const int32_t *keyv_f, const double *psi_c,
                                 const double *psi_f, double *x_c,
                                 double *x_f);
void oref_compress_and_accumulate_standard(
    int n1, int n2, int n3, const int nfl[3], const int nfu[3], int nseg_c,
    const int32_t *keyg_c, const int32_t *keyv_c, int nseg_f,
    const int32_t *keyg_f, const int32_t *keyv_f, const double *y_c,
    const double *y_f, double *psi_c, double *psi_f);

/* ---- potential (orbitals/potential.f90) ---- */
void oref_apply_potential_lr(int n1i, int n2i, int n3i, int n1ip, int n2ip,
                             int n3ip, const int ishift[3], double *psir,
                             const double *pot, double *epot);
/* with the confining term of confdata (NULL or potorder 0: none) */
void oref_apply_potential_lr_conf(int n1i, int n2i, int n3i, int n1ip, int n2ip,
                                  int n3ip, const int ishift[3], double *psir,
                                  const double *pot, const oref_confpot *cd,
                                  double *epot);

/* ---- free boundary conditions (growshrink.f90, convolut_ib_optim.f90) ---- */
void oref_comb_grow_all(int n1, int n2, int n3, const int nfl[3],
                        const int nfu[3], const double *x_c, const double *x_f,
                        double *psir);
void oref_comb_shrink(int n1, int n2, int n3, const int nfl[3],
                      const int nfu[3], const double *psir, double *y_c,
                      double *y_f);
void oref_convolkinetic_T(int n1, int n2, int n3, const int nfl[3],
                          const int nfu[3], const double hgrid[3],
                          const double *x_c, const double *x_f, double *y_c,
                          double *y_f, double *ekin);

/* ---- per-orbital drivers (modules/orbital_basis.f90:760-858) ---- */
/* periodic: psi/hpsi compressed; pot has n1i*n2i*n3i = (2n+2)^3 entries;
 * hpsi is accumulated (+=) */
void oref_locham_per(const int n[3], const double hgrid[3], int nseg_c,
                     int nvctr_c, const int32_t *keyg_c, const int32_t *keyv_c,
                     int nseg_f, int nvctr_f, const int32_t *keyg_f,
                     const int32_t *keyv_f, const double *pot,
                     const double *psi, double *hpsi, double *ekin,
                     double *epot);
/* free: pot has (2n+31)^3 entries laid out like psir */
void oref_locham_free(const int n[3], const int nfl[3], const int nfu[3],
                      const double hgrid[3], int nseg_c, int nvctr_c,
                      const int32_t *keyg_c, const int32_t *keyv_c,
                      int nseg_f, int nvctr_f, const int32_t *keyg_f,
                      const int32_t *keyv_f, const double *pot,
                      const double *psi, double *hpsi, double *ekin,
                      double *epot);
/* many orbitals sharing one locreg (the ket loop, hpsiortho.f90:945-958);
 * orbital loop serial, OpenMP inside the kernels */
void oref_locham_many(char geocode, const int n[3], const int nfl[3],
                      const int nfu[3], const double hgrid[3], int nseg_c,
                      int nvctr_c, const int32_t *keyg_c,
                      const int32_t *keyv_c, int nseg_f, int nvctr_f,
                      const int32_t *keyg_f, const int32_t *keyv_f,
                      const double *pot, int norb, const double *psi,
                      double *hpsi, double *ekin, double *epot);
/* the same with one confdata per orbital (NULL: none) */
void oref_locham_many_conf(char geocode, const int n[3], const int nfl[3],
                           const int nfu[3], const double hgrid[3], int nseg_c,
                           int nvctr_c, const int32_t *keyg_c,
                           const int32_t *keyv_c, int nseg_f, int nvctr_f,
                           const int32_t *keyg_f, const int32_t *keyv_f,
                           const double *pot, int norb, const double *psi,
                           double *hpsi, double *ekin, double *epot,
                           const oref_confpot *conf);

/* ---- density (orbitals/daubisf.f90:13-80, sumrho.f90:377-560) ---- */
void oref_daub_to_isf(char geocode, const int n[3], const int nfl[3],
                      const int nfu[3], int nseg_c, int nvctr_c,
                      const int32_t *keyg_c, const int32_t *keyv_c, int nseg_f,
                      int nvctr_f, const int32_t *keyg_f,
                      const int32_t *keyv_f, const double *psi, double *psir);
void oref_density(char geocode, const int n[3], const int nfl[3],
                  const int nfu[3], int nseg_c, int nvctr_c,
                  const int32_t *keyg_c, const int32_t *keyv_c, int nseg_f,
                  int nvctr_f, const int32_t *keyg_f, const int32_t *keyv_f,
                  int norb, const double *psi, const double *hfac,
                  const int *spin, double *rho);

/* ---- transfers (locregs/locreg_operations.f90:1665-1837, 2143-2380) ---- */
void oref_psi_to_locreg2(const int nl[3], const int ng[3], const int ns[3],
                         int nseg_lc, int nseg_lf, int nvctr_lc, int nvctr_lf,
                         const int32_t *keyg_l, const int32_t *keyv_l,
                         int nseg_gc, int nseg_gf, int nvctr_gc,
                         const int32_t *keyg_g, const int32_t *keyv_g,
                         const double *gpsi, double *lpsi);
int oref_lpsi_to_global2(const int nl[3], const int ng[3], const int ns[3],
                         int nseg_lc, int nseg_lf, int nvctr_lc, int nvctr_lf,
                         const int32_t *keyg_l, const int32_t *keyv_l,
                         int nseg_gc, int nseg_gf, int nvctr_gc,
                         const int32_t *keyg_g, const int32_t *keyv_g,
                         const double *lpsi, double *psi);

/* ---- descriptors (locregs/initlocregs.f90:1573-1686) ---- */
int oref_num_segkeys(int n1, int n2, int n3, int nl1, int nu1, int nl2,
                     int nu2, int nl3, int nu3, const uint8_t *logrid,
                     int *mseg, int *mvctr);
int oref_segkeys(int n1, int n2, int n3, int nl1, int nu1, int nl2, int nu2,
                 int nl3, int nu3, const uint8_t *logrid, int mseg,
                 int32_t *keyg, int32_t *keyv);

/* ---- convolution bounds (locregs/bounds.f90:555-1123) ----
 * The 17 integer arrays of convolutions_bounds in one int32 buffer, in the
 * order kb (ibyz_c ibxz_c ibxy_c ibyz_f ibxz_f ibxy_f), gb (ibzxx_c ibxxyy_c
 * ibyz_ff ibzxx_f ibxxyy_f), sb (ibzzx_c ibyyzz_c ibxy_ff ibzzx_f ibyyzz_f),
 * ibyyzz_r; each is the Fortran array ib(2, lo1:hi1, lo2:hi2) column-major.
 * layout: off[k] = first element of array k (off[17] = total), dims[4k..4k+3]
 * = lo1, hi1, lo2, hi2.  Returns nonzero when nfl > nfu on an axis. */
int oref_bounds_layout(const int n[3], const int nfl[3], const int nfu[3],
                       int64_t off[18], int32_t dims[68]);
int oref_locreg_bounds(const int n[3], const int nfl[3], const int nfu[3],
                       int nseg_c, const int32_t *keyg_c, int nseg_f,
                       const int32_t *keyg_f, int32_t *out);

/* ---- preconditioner, GPU part of preconditionall_OCL (interface_ocl.f90:
 * 622-880): ocl_preconditioner_generic_k (OpenCL/Preconditioner.c:245-358),
 * real case.  hgrid is the coarse spacing (the OpenCL code takes hgrid/2). */
void oref_precond_scal(const double hgrid[3], double c, double scal[8]);
void oref_precond_apply(char geocode, const int n[3], const double hgrid[3],
                        double c, int nseg_c, int nvctr_c, const int32_t *keyg_c,
                        const int32_t *keyv_c, int nseg_f, int nvctr_f,
                        const int32_t *keyg_f, const int32_t *keyv_f,
                        const double *x, double *y);
void oref_precond_cg(char geocode, const int n[3], const double hgrid[3], double c,
                     int ncong, int nseg_c, int nvctr_c, const int32_t *keyg_c,
                     const int32_t *keyv_c, int nseg_f, int nvctr_f,
                     const int32_t *keyg_f, const int32_t *keyv_f, double *x,
                     double *b);

int oref_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
