/*
 * CPU oracle for BigDFT's local-Hamiltonian path (Daubechies-16, FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Loaded by the
 * tests, by the smoke entry and by the CPU-baseline leg of the benchmark;
 * never by the product.
 *
 * Every routine restates a reference routine; the file:line it follows is
 * given in front of it (paths relative to the reference tree).  Storage is
 * Fortran column-major, keys are 1-based.
 *
 * Deliberate simplification (mathematically exact): the free-BC routines of
 * the reference (comb_grow_all, comb_shrink, ConvolkineticT,
 * apply_potential_lr_bounds) restrict their loops with "bounds" arrays that
 * only skip entries that are exactly zero.  The restatements here run over
 * the whole box instead; each output point that the reference computes gets
 * the same terms (plus exact zeros).  Summation order inside one output is
 * kept where it is cheap to keep.
 */
#include "bdft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Filter tables.                                                            */
/* ------------------------------------------------------------------------ */

/* Daubechies-16 low/high pass, index l = -7..8.
 * bigdft/src/orbitals/convolut_optim_per.f90:25-42 (ana_rot_per). */
static const double CH[16] = {
    -0.0033824159510050025955, -0.00054213233180001068935,
    0.031695087811525991431,   0.0076074873249766081919,
    -0.14329423835127266284,   -0.061273359067811077843,
    0.48135965125905339159,    0.77718575169962802862,
    0.36444189483617893676,    -0.051945838107881800736,
    -0.027219029917103486322,  0.049137179673730286787,
    0.0038087520138944894631,  -0.014952258337062199118,
    -0.00030292051472413308126, 0.0018899503327676891843};
static const double CG[16] = {
    -0.0018899503327676891843, -0.00030292051472413308126,
    0.014952258337062199118,   0.0038087520138944894631,
    -0.049137179673730286787,  -0.027219029917103486322,
    0.051945838107881800736,   0.36444189483617893676,
    -0.77718575169962802862,   0.48135965125905339159,
    0.061273359067811077843,   -0.14329423835127266284,
    -0.0076074873249766081919, 0.031695087811525991431,
    0.00054213233180001068935, -0.0033824159510050025955};
#define CHf(l) CH[(l) + 7]
#define CGf(l) CG[(l) + 7]

/* Magic filter, index l = -8..7.
 * bigdft/src/orbitals/convolut_optim_per.f90:537-553 (convrot_n_per). */
static const double MAGIC[16] = {
    8.4334247333529341094733325815816e-7,
    -0.1290557201342060969516786758559028e-4,
    0.8762984476210559564689161894116397e-4,
    -0.30158038132690463167163703826169879e-3,
    0.174723713672993903449447812749852942e-2,
    -0.942047030201080385922711540948195075e-2,
    0.2373821463724942397566389712597274535e-1,
    0.612625895831207982195380597e-1,
    0.9940415697834003993178616713,
    -0.604895289196983516002834636e-1,
    -0.2103025160930381434955489412839065067e-1,
    0.1337263414854794752733423467013220997e-1,
    -0.344128144493493857280881509686821861e-2,
    0.49443227688689919192282259476750972e-3,
    -0.5185986881173432922848639136911487e-4,
    2.72734492911979659657715313017228e-6};
#define MAGf(l) MAGIC[(l) + 8]

/* Second-derivative filter <phi|D^2|phi_l>, l = 0..14 (even in l).
 * bigdft/src/orbitals/convolut_optim_per.f90:1211-1225 and
 * bigdft/src/orbitals/convolut_ib_optim.f90:962-976 (same values). */
static const double KIN_A[15] = {
    -3.5536922899131901941296809374,
    2.2191465938911163898794546405,
    -0.6156141465570069496314853949,
    0.2371780582153805636239247476,
    -0.0822663999742123340987663521,
    0.02207029188482255523789911295638968409,
    -0.409765689342633823899327051188315485e-2,
    0.45167920287502235349480037639758496e-3,
    -0.2398228524507599670405555359023135e-4,
    2.0904234952920365957922889447361e-6,
    -3.7230763047369275848791496973044e-7,
    -1.05857055496741470373494132287e-8,
    -5.813879830282540547959250667e-11,
    2.70800493626319438269856689037647576e-13,
    -6.924474940639200152025730585882e-18};
/* <phi|D^2|psi_l>, l = -14..14.  convolut_ib_optim.f90:985-1013. */
static const double KIN_C[29] = {
    -3.869102413147656535541850057188e-18,
    1.5130616560866154733900029272077362e-13,
    -3.2264702314010525539061647271983988409e-11,
    -5.96264938781402337319841002642e-9,
    -2.1656830629214041470164889350342e-7,
    8.7969704055286288323596890609625e-7,
    -0.00001133456724516819987751818232711775,
    0.00021710795484646138591610188464622454,
    -0.0021356291838797986414312219042358542,
    0.00713761218453631422925717625758502986,
    -0.0284696165863973422636410524436931061,
    0.14327329352510759457155821037742893841,
    -0.42498050943780130143385739554118569733,
    0.65703074007121357894896358254040272157,
    -0.42081655293724308770919536332797729898,
    -0.21716117505137104371463587747283267899,
    0.63457035267892488185929915286969303251,
    -0.53298223962800395684936080758073568406,
    0.23370490631751294307619384973520033236,
    -0.05657736973328755112051544344507997075,
    0.0080872029411844780634067667008050127,
    -0.00093423623304808664741804536808932984,
    0.00005075807947289728306309081261461095,
    -4.62561497463184262755416490048242e-6,
    6.3919128513793415587294752371778e-7,
    1.87909235155149902916133888931e-8,
    1.04757345962781829480207861447155543883e-10,
    -4.84665690596158959648731537084025836e-13,
    1.2392629629188986192855777620877e-17};
/* <psi|D^2|psi_l>, l = 0..14 (even).  convolut_ib_optim.f90:1021-1038. */
static const double KIN_E[15] = {
    -24.875846029392331358907766562,
    -7.1440597663471719869313377994,
    -0.04251705323669172315864542163525830944,
    -0.26995931336279126953587091167128839196,
    0.08207454169225172612513390763444496516,
    -0.02207327034586634477996701627614752761,
    0.00409765642831595181639002667514310145,
    -0.00045167920287507774929432548999880117,
    0.00002398228524507599670405555359023135,
    -2.0904234952920365957922889447361e-6,
    3.7230763047369275848791496973044e-7,
    1.05857055496741470373494132287e-8,
    5.8138798302825405479592506674648873655e-11,
    -2.70800493626319438269856689037647576e-13,
    6.924474940639200152025730585882e-18};

/* Combined synthesis+magic filter fil2(j, s), j = -14..16, s = 1 (scaling),
 * 2 (wavelet).  bigdft/src/orbitals/v.inc:13-76 (identical in v_17.inc and
 * v_long.inc apart from zero padding). */
static const double FIL2[2][31] = {
    {-9.2250149921474082015e-9, 1.739330656197903949e-7,
     -1.5578172714074496288e-6, 9.7488426741565769756e-6,
     -0.000028480442793096733611, -0.000034162902320983755896,
     0.00054731040109973893777, -0.0034543155764822098308,
     -0.0043299995331208240697, 0.031835414345970668137,
     0.023424614735397134395, -0.13852477567405100085,
     -0.11023008167833446826, 0.41606693676621257536,
     0.78046048025074402849, 0.4259879826969471933,
     -0.014743648317973476936, -0.031312897214485659659,
     0.0437933264161550095, 0.0080208847559691655459,
     -0.01336723942112679464, -0.0017135145673762282241,
     0.0015547744749237938993, 0.00023987828928177507517,
     0.000025057266882581997398, -0.00001414752521104005874,
     2.0755675626515876328e-6, -4.0033729064957944938e-7,
     1.5691554225771085691e-7, -2.4646355857433108359e-8,
     1.5938753881171638248e-9},
    {-5.1545464563622292493e-9, 9.7186407588266778366e-8,
     -8.7796306212502899553e-7, 5.5890429568834202313e-6,
     -0.000017169835895815067013, -0.000011402461428743860613,
     0.00028479503466830096257, -0.0019699454319373247538,
     -0.0019602139595384833547, 0.015281781287179308495,
     0.0072888369266039604623, -0.039920047232423474172,
     -0.052935814515665507352, 0.049247868599583189247,
     0.40346722112554718347, -0.78110468414482695849,
     0.43810509543421050422, 0.077858113707232306275,
     -0.11950673101135450945, -0.022857677317437822508,
     0.028206423027424364667, 0.0037436631993987390831,
     -0.0027213277354099953613, -0.0004645045737819340081,
     -0.000040668222535397420851, 0.000024486891695262527278,
     -3.3186107540444071037e-6, 6.5211830898659553296e-7,
     -2.7666729880873517609e-7, 4.4109215856612868616e-8,
     -2.8525350339693075057e-9}};
#define FIL2f(j, s) FIL2[(s)-1][(j) + 14]

void oref_filters(double ch[16], double cg[16], double magic[16],
                  double kin_a[15], double kin_c[29], double kin_e[15],
                  double fil2[62]) {
  memcpy(ch, CH, sizeof CH);
  memcpy(cg, CG, sizeof CG);
  memcpy(magic, MAGIC, sizeof MAGIC);
  memcpy(kin_a, KIN_A, sizeof KIN_A);
  memcpy(kin_c, KIN_C, sizeof KIN_C);
  memcpy(kin_e, KIN_E, sizeof KIN_E);
  memcpy(fil2, FIL2, sizeof FIL2);
}

int oref_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* fill_mod_arr, bigdft/src/orbitals/convolut_optim_per.f90:1600-1629 */
static inline int floordiv2(int a) { return a >= 0 ? a / 2 : -((-a + 1) / 2); }

static inline int pmod(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

/* ------------------------------------------------------------------------ */
/* Periodic rotating 1D kernels.                                             */
/* ------------------------------------------------------------------------ */

/* ana_rot_per, convolut_optim_per.f90:10-144.
 * x(0:2n+1, ndat) -> y(ndat, 0:2n+1), c in y(:,0:n), d in y(:,n+1:2n+1). */
void oref_ana_rot_per(int n, int ndat, const double *x, double *y) {
  const int m = 2 * n + 2;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ndat; ++j) {
    const double *xj = x + (size_t)m * j;
    for (int i = 0; i <= n; ++i) {
      double ci = 0.0, di = 0.0;
      for (int l = -7; l <= 8; ++l) {
        const int k = pmod(l + 2 * i, m);
        ci += CHf(l) * xj[k];
        di += CGf(l) * xj[k];
      }
      y[j + (size_t)ndat * i] = ci;
      y[j + (size_t)ndat * (n + 1 + i)] = di;
    }
  }
}

/* syn_rot_per, convolut_optim_per.f90:148-304.
 * x(0:2n+1, ndat) with c in 0:n and d in n+1:2n+1 -> y(ndat, 0:2n+1). */
void oref_syn_rot_per(int n, int ndat, const double *x, double *y) {
  const int m = 2 * n + 2, np1 = n + 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ndat; ++j) {
    const double *xj = x + (size_t)m * j;
    for (int i = 0; i <= n; ++i) {
      int k = pmod(i - 4, np1);
      double se = CHf(8) * xj[k] + CGf(8) * xj[np1 + k];
      k = pmod(i + 4, np1);
      double so = CHf(-7) * xj[k] + CGf(-7) * xj[np1 + k];
      for (int l = -3; l <= 3; ++l) {
        k = pmod(i - l, np1);
        se = se + CHf(2 * l) * xj[k] + CGf(2 * l) * xj[np1 + k];
        so = so + CHf(2 * l + 1) * xj[k] + CGf(2 * l + 1) * xj[np1 + k];
      }
      y[j + (size_t)ndat * (2 * i)] = se;
      y[j + (size_t)ndat * (2 * i + 1)] = so;
    }
  }
}

/* convrot_n_per, convolut_optim_per.f90:519-643: y(j,i)=sum_l fil(l) x(i+l) */
void oref_convrot_n_per(int n1, int ndat, const double *x, double *y) {
  const int m = n1 + 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ndat; ++j) {
    const double *xj = x + (size_t)m * j;
    for (int i = 0; i <= n1; ++i) {
      double tt = 0.0;
      for (int l = -8; l <= 7; ++l) tt += xj[pmod(i + l, m)] * MAGf(l);
      y[j + (size_t)ndat * i] = tt;
    }
  }
}

/* convrot_t_per, convolut_optim_per.f90:646-766: reversed taps l=-7..8 */
void oref_convrot_t_per(int n1, int ndat, const double *x, double *y) {
  const int m = n1 + 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ndat; ++j) {
    const double *xj = x + (size_t)m * j;
    for (int i = 0; i <= n1; ++i) {
      double tt = 0.0;
      for (int l = -7; l <= 8; ++l) tt += xj[pmod(i + l, m)] * MAGf(-l);
      y[j + (size_t)ndat * i] = tt;
    }
  }
}

/* convolut_kinetic_per_T + conv_kin_x/y/z,
 * convolut_optim_per.f90:1169-1260, 1372-1597.
 * y += -1/2 Laplacian x along each axis; kstrten(d) = sum x * (K_d x). */
void oref_convolut_kinetic_per_T(int n1, int n2, int n3, const double hgrid[3],
                                 const double *x, double *y,
                                 double kstrten[3]) {
  double fil[3][29];
  for (int d = 0; d < 3; ++d) {
    const double scale = -0.5 / (hgrid[d] * hgrid[d]);
    for (int l = 0; l <= 14; ++l) {
      fil[d][14 + l] = KIN_A[l] * scale;
      fil[d][14 - l] = KIN_A[l] * scale;
    }
  }
  const int m1 = n1 + 1, m2 = n2 + 1, m3 = n3 + 1;
  const size_t s2 = (size_t)m1, s3 = (size_t)m1 * m2;
  double e1 = 0.0, e2 = 0.0, e3 = 0.0;
  /* x direction */
#pragma omp parallel for reduction(+ : e1) schedule(static)
  for (int i3 = 0; i3 < m3; ++i3)
    for (int i2 = 0; i2 < m2; ++i2) {
      const size_t b = i2 * s2 + i3 * s3;
      for (int i1 = 0; i1 < m1; ++i1) {
        double tt = 0.0;
        for (int l = -14; l <= 14; ++l)
          tt += x[b + pmod(i1 + l, m1)] * fil[0][l + 14];
        y[b + i1] += tt;
        e1 += tt * x[b + i1];
      }
    }
  /* y direction */
#pragma omp parallel for reduction(+ : e2) schedule(static)
  for (int i3 = 0; i3 < m3; ++i3)
    for (int i2 = 0; i2 < m2; ++i2)
      for (int i1 = 0; i1 < m1; ++i1) {
        const size_t b = i1 + i3 * s3;
        double tt = 0.0;
        for (int l = -14; l <= 14; ++l)
          tt += x[b + pmod(i2 + l, m2) * s2] * fil[1][l + 14];
        y[b + i2 * s2] += tt;
        e2 += tt * x[b + i2 * s2];
      }
  /* z direction */
#pragma omp parallel for reduction(+ : e3) schedule(static)
  for (int i3 = 0; i3 < m3; ++i3)
    for (int i2 = 0; i2 < m2; ++i2)
      for (int i1 = 0; i1 < m1; ++i1) {
        const size_t b = i1 + i2 * s2;
        double tt = 0.0;
        for (int l = -14; l <= 14; ++l)
          tt += x[b + pmod(i3 + l, m3) * s3] * fil[2][l + 14];
        y[b + i3 * s3] += tt;
        e3 += tt * x[b + i3 * s3];
      }
  kstrten[0] = e1;
  kstrten[1] = e2;
  kstrten[2] = e3;
}

/* ------------------------------------------------------------------------ */
/* Periodic 3D drivers, convolut_common_per.f90.                             */
/* ------------------------------------------------------------------------ */

/* synthese_per_self, convolut_common_per.f90:87-109 (x is overwritten) */
void oref_synthese_per_self(int n1, int n2, int n3, double *x, double *y) {
  oref_syn_rot_per(n1, (2 * n2 + 2) * (2 * n3 + 2), x, y);
  oref_syn_rot_per(n2, (2 * n3 + 2) * (2 * n1 + 2), y, x);
  oref_syn_rot_per(n3, (2 * n1 + 2) * (2 * n2 + 2), x, y);
}

/* analyse_per_self, convolut_common_per.f90:38-57 (y is overwritten) */
void oref_analyse_per_self(int n1, int n2, int n3, double *y, double *x) {
  oref_ana_rot_per(n1, (2 * n2 + 2) * (2 * n3 + 2), y, x);
  oref_ana_rot_per(n2, (2 * n3 + 2) * (2 * n1 + 2), x, y);
  oref_ana_rot_per(n3, (2 * n1 + 2) * (2 * n2 + 2), y, x);
}

/* convolut_magic_n_per, convolut_common_per.f90:112-137 (n = fine max idx) */
void oref_convolut_magic_n_per(int n1, int n2, int n3, const double *x,
                               double *y, double *ww) {
  oref_convrot_n_per(n1, (n2 + 1) * (n3 + 1), x, y);
  oref_convrot_n_per(n2, (n3 + 1) * (n1 + 1), y, ww);
  oref_convrot_n_per(n3, (n1 + 1) * (n2 + 1), ww, y);
}

/* convolut_magic_t_per_self, convolut_common_per.f90:166-190 */
void oref_convolut_magic_t_per_self(int n1, int n2, int n3, double *x,
                                    double *y) {
  oref_convrot_t_per(n1, (n2 + 1) * (n3 + 1), x, y);
  oref_convrot_t_per(n2, (n3 + 1) * (n1 + 1), y, x);
  oref_convrot_t_per(n3, (n1 + 1) * (n2 + 1), x, y);
}

/* ------------------------------------------------------------------------ */
/* Compression, wavelib/un-compress.f90.                                     */
/* ------------------------------------------------------------------------ */

/* segment decode shared by every (un)compress routine, e.g.
 * un-compress.f90:903-913: keyg(1) = 1-based linear start i3*(n1+1)(n2+1) +
 * i2*(n1+1) + i1 + 1, keyg(2) = end, keyv = 1-based vector offset. */
static inline void seg_decode(int n1, int n2, const int32_t *keyg,
                              const int32_t *keyv, int iseg, int *jj, int *i0,
                              int *i1, int *i2, int *i3) {
  const int j0 = keyg[2 * iseg], j1 = keyg[2 * iseg + 1];
  int ii = j0 - 1;
  *i3 = ii / ((n1 + 1) * (n2 + 1));
  ii -= *i3 * (n1 + 1) * (n2 + 1);
  *i2 = ii / (n1 + 1);
  *i0 = ii - *i2 * (n1 + 1);
  *i1 = *i0 + j1 - j0;
  *jj = keyv[iseg];
}

/* psig(0:n1,2,0:n2,2,0:n3,2) index */
static inline size_t PSIG(int n1, int n2, int n3, int i1, int s1, int i2,
                          int s2, int i3, int s3) {
  const size_t m1 = (size_t)(n1 + 1), m2 = (size_t)(n2 + 1),
               m3 = (size_t)(n3 + 1);
  (void)m3;
  return ((size_t)i1 + m1 * s1) +
         2 * m1 * (((size_t)i2 + m2 * s2) + 2 * m2 * ((size_t)i3 + m3 * s3));
}

/* fine component t = 1..7 <-> (s1,s2,s3): un-compress.f90:930-936 */
static const int FSEL[7][3] = {{1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1},
                               {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

/* uncompress_per, un-compress.f90:877-948, without the final synthesis */
void oref_uncompress_psig(int n1, int n2, int n3, int nseg_c,
                          const int32_t *keyg_c, const int32_t *keyv_c,
                          int nseg_f, const int32_t *keyg_f,
                          const int32_t *keyv_f, const double *psi_c,
                          const double *psi_f, double *psig) {
  const size_t tot = (size_t)8 * (n1 + 1) * (n2 + 1) * (n3 + 1);
  memset(psig, 0, tot * sizeof(double));
  for (int iseg = 0; iseg < nseg_c; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_c, keyv_c, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      psig[PSIG(n1, n2, n3, i, 0, i2, 0, i3, 0)] = psi_c[i - i0 + jj - 1];
  }
  for (int iseg = 0; iseg < nseg_f; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_f, keyv_f, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      for (int t = 0; t < 7; ++t)
        psig[PSIG(n1, n2, n3, i, FSEL[t][0], i2, FSEL[t][1], i3,
                  FSEL[t][2])] = psi_f[7 * (size_t)(i - i0 + jj - 1) + t];
  }
}

void oref_uncompress_per(int n1, int n2, int n3, int nseg_c, int nvctr_c,
                         const int32_t *keyg_c, const int32_t *keyv_c,
                         int nseg_f, int nvctr_f, const int32_t *keyg_f,
                         const int32_t *keyv_f, const double *psi_c,
                         const double *psi_f, double *psifscf, double *psig) {
  (void)nvctr_c;
  (void)nvctr_f;
  oref_uncompress_psig(n1, n2, n3, nseg_c, keyg_c, keyv_c, nseg_f, keyg_f,
                       keyv_f, psi_c, psi_f, psig);
  oref_synthese_per_self(n1, n2, n3, psig, psifscf);
}

/* compress_and_accumulate_mixed, un-compress.f90:2174-2238 */
void oref_compress_and_accumulate_mixed(int n1, int n2, int n3, int nseg_c,
                                        const int32_t *keyg_c,
                                        const int32_t *keyv_c, int nseg_f,
                                        const int32_t *keyg_f,
                                        const int32_t *keyv_f,
                                        const double *psig, double *psi_c,
                                        double *psi_f) {
  for (int iseg = 0; iseg < nseg_c; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_c, keyv_c, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      psi_c[i - i0 + jj - 1] += psig[PSIG(n1, n2, n3, i, 0, i2, 0, i3, 0)];
  }
  for (int iseg = 0; iseg < nseg_f; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_f, keyv_f, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      for (int t = 0; t < 7; ++t)
        psi_f[7 * (size_t)(i - i0 + jj - 1) + t] +=
            psig[PSIG(n1, n2, n3, i, FSEL[t][0], i2, FSEL[t][1], i3,
                      FSEL[t][2])];
  }
}

/* x_f(7, nfl1:nfu1, nfl2:nfu2, nfl3:nfu3) index */
static inline size_t XF(const int nfl[3], const int nfu[3], int t, int i1,
                        int i2, int i3) {
  const size_t m1 = (size_t)(nfu[0] - nfl[0] + 1),
               m2 = (size_t)(nfu[1] - nfl[1] + 1);
  return (size_t)t + 7 * ((size_t)(i1 - nfl[0]) +
                          m1 * ((size_t)(i2 - nfl[1]) + m2 * (size_t)(i3 - nfl[2])));
}

/* uncompress_forstandard, un-compress.f90:286-360 (scal = 1, the transposed
 * copies x_f1..x_f3 are not needed by the dense kinetic below) */
void oref_uncompress_forstandard(int n1, int n2, int n3, const int nfl[3],
                                 const int nfu[3], int nseg_c,
                                 const int32_t *keyg_c, const int32_t *keyv_c,
                                 int nseg_f, const int32_t *keyg_f,
                                 const int32_t *keyv_f, const double *psi_c,
                                 const double *psi_f, double *x_c,
                                 double *x_f) {
  const size_t m1 = (size_t)(n1 + 1), m2 = (size_t)(n2 + 1);
  memset(x_c, 0, sizeof(double) * m1 * m2 * (size_t)(n3 + 1));
  memset(x_f, 0,
         sizeof(double) * 7 * (size_t)(nfu[0] - nfl[0] + 1) *
             (size_t)(nfu[1] - nfl[1] + 1) * (size_t)(nfu[2] - nfl[2] + 1));
  for (int iseg = 0; iseg < nseg_c; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_c, keyv_c, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      x_c[i + m1 * (i2 + m2 * i3)] = psi_c[i - i0 + jj - 1];
  }
  for (int iseg = 0; iseg < nseg_f; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_f, keyv_f, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      for (int t = 0; t < 7; ++t)
        x_f[XF(nfl, nfu, t, i, i2, i3)] = psi_f[7 * (size_t)(i - i0 + jj - 1) + t];
  }
}

/* compress_and_accumulate_standard, un-compress.f90:2105-2170 */
void oref_compress_and_accumulate_standard(
    int n1, int n2, int n3, const int nfl[3], const int nfu[3], int nseg_c,
    const int32_t *keyg_c, const int32_t *keyv_c, int nseg_f,
    const int32_t *keyg_f, const int32_t *keyv_f, const double *y_c,
    const double *y_f, double *psi_c, double *psi_f) {
  const size_t m1 = (size_t)(n1 + 1), m2 = (size_t)(n2 + 1);
  (void)n3;
  for (int iseg = 0; iseg < nseg_c; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_c, keyv_c, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      psi_c[i - i0 + jj - 1] += y_c[i + m1 * (i2 + m2 * i3)];
  }
  for (int iseg = 0; iseg < nseg_f; ++iseg) {
    int jj, i0, i1, i2, i3;
    seg_decode(n1, n2, keyg_f, keyv_f, iseg, &jj, &i0, &i1, &i2, &i3);
    for (int i = i0; i <= i1; ++i)
      for (int t = 0; t < 7; ++t)
        psi_f[7 * (size_t)(i - i0 + jj - 1) + t] += y_f[XF(nfl, nfu, t, i, i2, i3)];
  }
}

/* ------------------------------------------------------------------------ */
/* Potential, orbitals/potential.f90.                                        */
/* ------------------------------------------------------------------------ */

/* apply_potential_lr_nobounds, potential.f90:850-1045 (nspinor = 1 loop
 * :1011-1035).  psir outside the potential box is zeroed (:874-947). */
void oref_apply_potential_lr(int n1i, int n2i, int n3i, int n1ip, int n2ip,
                             int n3ip, const int ishift[3], double *psir,
                             const double *pot, double *epot) {
  const int i3s = ishift[2] + 1 > 1 ? ishift[2] + 1 : 1;
  const int i3e = n3ip + ishift[2] < n3i ? n3ip + ishift[2] : n3i;
  const int i2s = ishift[1] + 1 > 1 ? ishift[1] + 1 : 1;
  const int i2e = n2ip + ishift[1] < n2i ? n2ip + ishift[1] : n2i;
  const int i1s = ishift[0] + 1 > 1 ? ishift[0] + 1 : 1;
  const int i1e = n1ip + ishift[0] < n1i ? n1ip + ishift[0] : n1i;
  double ep = 0.0;
#pragma omp parallel for reduction(+ : ep) schedule(static)
  for (int i3 = 1; i3 <= n3i; ++i3)
    for (int i2 = 1; i2 <= n2i; ++i2) {
      double *row = psir + (size_t)n1i * ((i2 - 1) + (size_t)n2i * (i3 - 1));
      const int in23 = (i3 >= i3s && i3 <= i3e && i2 >= i2s && i2 <= i2e);
      for (int i1 = 1; i1 <= n1i; ++i1) {
        if (in23 && i1 >= i1s && i1 <= i1e) {
          const double psir1 = row[i1 - 1];
          const double pot1 =
              pot[(i1 - ishift[0] - 1) +
                  (size_t)n1ip * ((i2 - ishift[1] - 1) +
                                  (size_t)n2ip * (i3 - ishift[2] - 1))];
          const double tt11 = pot1 * psir1;
          ep += tt11 * psir1;
          row[i1 - 1] = tt11;
        } else {
          row[i1 - 1] = 0.0;
        }
      }
    }
  *epot = ep;
}

/* apply_potential_lr_conf_nobounds (potential.f90:362-560), nspinor = 1:
 * as oref_apply_potential_lr with the confining term
 * prefac * r2^(potorder/2), r2 = sum_d (hh_d (i_d + ioffset_d) - rxyzConf_d)^2
 * (1-based ISF index i_d of the wavefunction box), added to the potential
 * where the potential box and the wavefunction box overlap (:541-556).  The
 * free-BC variant apply_potential_lr_conf (:203-360) restricts the product
 * to the ibyyzz_r rows, outside of which psir is zero: same result. */
void oref_apply_potential_lr_conf(int n1i, int n2i, int n3i, int n1ip, int n2ip,
                                  int n3ip, const int ishift[3], double *psir,
                                  const double *pot, const oref_confpot *cd,
                                  double *epot) {
  if (!cd || cd->potorder == 0) {
    oref_apply_potential_lr(n1i, n2i, n3i, n1ip, n2ip, n3ip, ishift, psir, pot, epot);
    return;
  }
  const int i3s = ishift[2] + 1 > 1 ? ishift[2] + 1 : 1;
  const int i3e = n3ip + ishift[2] < n3i ? n3ip + ishift[2] : n3i;
  const int i2s = ishift[1] + 1 > 1 ? ishift[1] + 1 : 1;
  const int i2e = n2ip + ishift[1] < n2i ? n2ip + ishift[1] : n2i;
  const int i1s = ishift[0] + 1 > 1 ? ishift[0] + 1 : 1;
  const int i1e = n1ip + ishift[0] < n1i ? n1ip + ishift[0] : n1i;
  const int half = cd->potorder / 2;
  double ep = 0.0;
#pragma omp parallel for reduction(+ : ep) schedule(static)
  for (int i3 = 1; i3 <= n3i; ++i3) {
    const double z = cd->hh[2] * (double)(i3 + cd->ioffset[2]) - cd->rxyzConf[2];
    const double z2 = z * z;
    for (int i2 = 1; i2 <= n2i; ++i2) {
      const double y = cd->hh[1] * (double)(i2 + cd->ioffset[1]) - cd->rxyzConf[1];
      const double y2 = y * y;
      double *row = psir + (size_t)n1i * ((i2 - 1) + (size_t)n2i * (i3 - 1));
      const int in23 = (i3 >= i3s && i3 <= i3e && i2 >= i2s && i2 <= i2e);
      for (int i1 = 1; i1 <= n1i; ++i1) {
        if (in23 && i1 >= i1s && i1 <= i1e) {
          const double x = cd->hh[0] * (double)(i1 + cd->ioffset[0]) - cd->rxyzConf[0];
          const double r2 = x * x + y2 + z2;
          double rp = 1.0;
          for (int k = 0; k < half; ++k) rp *= r2;
          const double cp = cd->prefac * rp;
          const double psir1 = row[i1 - 1];
          const double pot1 =
              pot[(i1 - ishift[0] - 1) +
                  (size_t)n1ip * ((i2 - ishift[1] - 1) +
                                  (size_t)n2ip * (i3 - ishift[2] - 1))] +
              cp;
          const double tt11 = pot1 * psir1;
          ep += tt11 * psir1;
          row[i1 - 1] = tt11;
        } else {
          row[i1 - 1] = 0.0;
        }
      }
    }
  }
  *epot = ep;
}

/* ------------------------------------------------------------------------ */
/* Free boundary conditions.                                                 */
/* ------------------------------------------------------------------------ */

/* One rotating grow pass: x(0:2n+1, ndat) with scaling coefficients in 0:n
 * and wavelet coefficients in n+1:2n+1 -> y(ndat, -14:2n+16), using
 * y(2i+j) = sum fil2(j,1) c(i) + fil2(j,2) d(i).  This is the 1D operator of
 * comb_rot_grow_loc_1/_2/_square (growshrink.f90:116-308,
 * combined_grow_optim.f90:16-513), applied to both the coarse part and the
 * tree part at once. */
static void grow_rot(int n, int ndat, const double *x, double *y) {
  const int m = 2 * n + 2, nout = 2 * n + 31;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ndat; ++j) {
    const double *xj = x + (size_t)m * j;
    for (int p = 0; p < nout; ++p) {
      const int I = p - 14;
      double tt = 0.0;
      int lo = -floordiv2(16 - I); /* ceil((I-16)/2) */
      int hi = floordiv2(I + 14);
      if (lo < 0) lo = 0;
      if (hi > n) hi = n;
      for (int i = lo; i <= hi; ++i) {
        const int jj = I - 2 * i;
        tt += FIL2f(jj, 1) * xj[i] + FIL2f(jj, 2) * xj[n + 1 + i];
      }
      y[j + (size_t)ndat * p] = tt;
    }
  }
}

/* One rotating shrink pass: x(-14:2n+16, ndat) -> y(ndat, 0:2n+1):
 * c(i) = sum_j fil2(j,1) x(2i+j), d(i) = sum_j fil2(j,2) x(2i+j), as in
 * comb_rot_shrink_loc(_1/_2/_3) (combined_shrink_optim.f90:14-299,
 * growshrink.f90:419-473). */
static void shrink_rot(int n, int ndat, const double *x, double *y) {
  const int m = 2 * n + 31;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ndat; ++j) {
    const double *xj = x + (size_t)m * j;
    for (int i = 0; i <= n; ++i) {
      double ci = 0.0, di = 0.0;
      for (int l = -14; l <= 16; ++l) {
        const double v = xj[2 * i + l + 14];
        ci += FIL2f(l, 1) * v;
        di += FIL2f(l, 2) * v;
      }
      y[j + (size_t)ndat * i] = ci;
      y[j + (size_t)ndat * (n + 1 + i)] = di;
    }
  }
}

/* Build the mixed layout psig(0:n1,2,0:n2,2,0:n3,2) from the standard
 * coarse/fine arrays (zero outside the fine box). */
static void standard_to_psig(int n1, int n2, int n3, const int nfl[3],
                             const int nfu[3], const double *x_c,
                             const double *x_f, double *psig) {
  const size_t m1 = (size_t)(n1 + 1), m2 = (size_t)(n2 + 1);
  memset(psig, 0, sizeof(double) * 8 * m1 * m2 * (size_t)(n3 + 1));
  for (int i3 = 0; i3 <= n3; ++i3)
    for (int i2 = 0; i2 <= n2; ++i2)
      for (int i1 = 0; i1 <= n1; ++i1)
        psig[PSIG(n1, n2, n3, i1, 0, i2, 0, i3, 0)] = x_c[i1 + m1 * (i2 + m2 * i3)];
  for (int i3 = nfl[2]; i3 <= nfu[2]; ++i3)
    for (int i2 = nfl[1]; i2 <= nfu[1]; ++i2)
      for (int i1 = nfl[0]; i1 <= nfu[0]; ++i1)
        for (int t = 0; t < 7; ++t)
          psig[PSIG(n1, n2, n3, i1, FSEL[t][0], i2, FSEL[t][1], i3, FSEL[t][2])] =
              x_f[XF(nfl, nfu, t, i1, i2, i3)];
}

/* comb_grow_all, growshrink.f90:12-37 (coarse part comb_grow_c :49-76 plus
 * tree part comb_grow_tree :78-113): psir(-14:2n1+16, -14:2n2+16, -14:2n3+16)
 * = combined synthesis+magic of the coarse and fine coefficients. */
void oref_comb_grow_all(int n1, int n2, int n3, const int nfl[3],
                        const int nfu[3], const double *x_c, const double *x_f,
                        double *psir) {
  const size_t w0 = (size_t)8 * (n1 + 1) * (n2 + 1) * (n3 + 1);
  const size_t big = (size_t)(2 * n1 + 31) * (2 * n2 + 31) * (2 * n3 + 31);
  double *psig = malloc(sizeof(double) * w0);
  double *w1 = malloc(sizeof(double) * (big > w0 * 4 ? big : w0 * 4));
  double *w2 = malloc(sizeof(double) * (big > w0 * 4 ? big : w0 * 4));
  standard_to_psig(n1, n2, n3, nfl, nfu, x_c, x_f, psig);
  /* i1,i2,i3 -> i2,i3,I1 -> i3,I1,I2 -> I1,I2,I3 */
  grow_rot(n1, (2 * n2 + 2) * (2 * n3 + 2), psig, w1);
  grow_rot(n2, (2 * n3 + 2) * (2 * n1 + 31), w1, w2);
  grow_rot(n3, (2 * n1 + 31) * (2 * n2 + 31), w2, psir);
  free(psig);
  free(w1);
  free(w2);
}

/* comb_shrink, growshrink.f90:310-340: combined magic^T + analysis back to
 * y_c(0:n1,0:n2,0:n3) and y_f(7,nfl1:nfu1,nfl2:nfu2,nfl3:nfu3). */
void oref_comb_shrink(int n1, int n2, int n3, const int nfl[3],
                      const int nfu[3], const double *psir, double *y_c,
                      double *y_f) {
  const size_t w0 = (size_t)8 * (n1 + 1) * (n2 + 1) * (n3 + 1);
  const size_t big = (size_t)(2 * n1 + 31) * (2 * n2 + 31) * (2 * n3 + 31);
  double *w1 = malloc(sizeof(double) * big);
  double *w2 = malloc(sizeof(double) * big);
  double *psig = malloc(sizeof(double) * w0);
  /* I1,I2,I3 -> I2,I3,i1 -> I3,i1,i2 -> i1,i2,i3 */
  shrink_rot(n1, (2 * n2 + 31) * (2 * n3 + 31), psir, w1);
  shrink_rot(n2, (2 * n3 + 31) * (2 * n1 + 2), w1, w2);
  shrink_rot(n3, (2 * n1 + 2) * (2 * n2 + 2), w2, psig);
  const size_t m1 = (size_t)(n1 + 1), m2 = (size_t)(n2 + 1);
  for (int i3 = 0; i3 <= n3; ++i3)
    for (int i2 = 0; i2 <= n2; ++i2)
      for (int i1 = 0; i1 <= n1; ++i1)
        y_c[i1 + m1 * (i2 + m2 * i3)] = psig[PSIG(n1, n2, n3, i1, 0, i2, 0, i3, 0)];
  for (int i3 = nfl[2]; i3 <= nfu[2]; ++i3)
    for (int i2 = nfl[1]; i2 <= nfu[1]; ++i2)
      for (int i1 = nfl[0]; i1 <= nfu[0]; ++i1)
        for (int t = 0; t < 7; ++t)
          y_f[XF(nfl, nfu, t, i1, i2, i3)] =
              psig[PSIG(n1, n2, n3, i1, FSEL[t][0], i2, FSEL[t][1], i3, FSEL[t][2])];
  free(w1);
  free(w2);
  free(psig);
}

/* ConvolkineticT, convolut_ib_optim.f90:921-1740: kinetic operator in the
 * wavelet basis with the 2x2 filter block per axis,
 *   c_out(i) += sum_l a(l) c(i+l) + b(l) d(i+l)
 *   d_out(i) += sum_l c(l) c(i+l) + e(l) d(i+l),   b(l) = c(-l),
 * all filters scaled by -1/2/h_d^2 with the coarse grid spacing, zero outside
 * the box (free BC).  ekin = sum of (added term) * x at each output point.
 * y_c, y_f are accumulated. */
void oref_convolkinetic_T(int n1, int n2, int n3, const int nfl[3],
                          const int nfu[3], const double hgrid[3],
                          const double *x_c, const double *x_f, double *y_c,
                          double *y_f, double *ekin) {
  const int n[3] = {n1, n2, n3};
  const size_t w0 = (size_t)8 * (n1 + 1) * (n2 + 1) * (n3 + 1);
  double *xg = malloc(sizeof(double) * w0);
  double *yg = calloc(w0, sizeof(double));
  standard_to_psig(n1, n2, n3, nfl, nfu, x_c, x_f, xg);
  double F[3][2][2][29];
  for (int d = 0; d < 3; ++d) {
    const double scale = -0.5 / (hgrid[d] * hgrid[d]);
    for (int l = -14; l <= 14; ++l) {
      const int al = l < 0 ? -l : l;
      F[d][0][0][l + 14] = scale * KIN_A[al];
      F[d][0][1][l + 14] = scale * KIN_C[-l + 14]; /* b(l) = c(-l) */
      F[d][1][0][l + 14] = scale * KIN_C[l + 14];
      F[d][1][1][l + 14] = scale * KIN_E[al];
    }
  }
  const size_t W1 = 2 * (size_t)(n1 + 1), W2 = 2 * (size_t)(n2 + 1),
               W3 = 2 * (size_t)(n3 + 1);
  const size_t stride[3] = {1, W1, W1 * W2};
  const size_t W[3] = {W1, W2, W3};
  double ek = 0.0;
  for (int d = 0; d < 3; ++d) {
    const int m = n[d] + 1;
    /* the two other axes */
    const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
    double ekd = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : ekd) schedule(static)
    for (size_t j2 = 0; j2 < W[a2]; ++j2)
      for (size_t j1 = 0; j1 < W[a1]; ++j1) {
        const size_t base = j1 * stride[a1] + j2 * stride[a2];
        for (int so = 0; so < 2; ++so)
          for (int i = 0; i < m; ++i) {
            double tt = 0.0;
            for (int l = -14; l <= 14; ++l) {
              const int t = i + l;
              if (t < 0 || t >= m) continue;
              tt += xg[base + stride[d] * (size_t)t] * F[d][so][0][l + 14] +
                    xg[base + stride[d] * (size_t)(m + t)] * F[d][so][1][l + 14];
            }
            const size_t o = base + stride[d] * (size_t)(so * m + i);
            yg[o] += tt;
            ekd += tt * xg[o];
          }
      }
    ek += ekd;
  }
  /* accumulate into y_c, y_f */
  const size_t m1 = (size_t)(n1 + 1), m2 = (size_t)(n2 + 1);
  for (int i3 = 0; i3 <= n3; ++i3)
    for (int i2 = 0; i2 <= n2; ++i2)
      for (int i1 = 0; i1 <= n1; ++i1)
        y_c[i1 + m1 * (i2 + m2 * i3)] += yg[PSIG(n1, n2, n3, i1, 0, i2, 0, i3, 0)];
  for (int i3 = nfl[2]; i3 <= nfu[2]; ++i3)
    for (int i2 = nfl[1]; i2 <= nfu[1]; ++i2)
      for (int i1 = nfl[0]; i1 <= nfu[0]; ++i1)
        for (int t = 0; t < 7; ++t)
          y_f[XF(nfl, nfu, t, i1, i2, i3)] +=
              yg[PSIG(n1, n2, n3, i1, FSEL[t][0], i2, FSEL[t][1], i3, FSEL[t][2])];
  *ekin = ek;
  free(xg);
  free(yg);
}

/* ------------------------------------------------------------------------ */
/* Per-orbital drivers: local_hamiltonian_ket, orbital_basis.f90:760-858.    */
/* ------------------------------------------------------------------------ */

/* periodic branch: daub_to_isf_locham (daubisf.f90:297-312) ->
 * psir_to_vpsi (locreg_operations.f90:2502-2590) -> isf_to_daub_kinetic P
 * (locreg_operations.f90:2839-2865) */
static void locham_per(const int n[3], const double hgrid[3], int nseg_c,
                       int nvctr_c, const int32_t *keyg_c, const int32_t *keyv_c,
                       int nseg_f, int nvctr_f, const int32_t *keyg_f,
                       const int32_t *keyv_f, const double *pot,
                       const double *psi, double *hpsi, double *ekin,
                       double *epot, const oref_confpot *cd) {
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const int n1i = 2 * n1 + 2, n2i = 2 * n2 + 2, n3i = 2 * n3 + 2;
  const size_t tot = (size_t)n1i * n2i * n3i;
  double *x_c = malloc(sizeof(double) * tot);
  double *y_c = malloc(sizeof(double) * tot);
  double *psir = malloc(sizeof(double) * tot);
  const double *psi_f = nvctr_f > 0 ? psi + nvctr_c : psi;
  double *hpsi_f = nvctr_f > 0 ? hpsi + nvctr_c : hpsi;
  /* daub_to_isf_locham: uncompress_per (psig in psir, fine scaling in x_c)
   * then convolut_magic_n_per(x_c -> psir, work y_c) */
  oref_uncompress_per(n1, n2, n3, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f,
                      nvctr_f, keyg_f, keyv_f, psi, psi_f, x_c, psir);
  oref_convolut_magic_n_per(n1i - 1, n2i - 1, n3i - 1, x_c, psir, y_c);
  /* psir_to_vpsi: apply_potential_lr_nobounds, ishift = 0 */
  const int ishift[3] = {0, 0, 0};
  oref_apply_potential_lr_conf(n1i, n2i, n3i, n1i, n2i, n3i, ishift, psir, pot,
                               cd, epot);
  /* isf_to_daub_kinetic, periodic: hgridh = hgrid/2 */
  const double hh[3] = {0.5 * hgrid[0], 0.5 * hgrid[1], 0.5 * hgrid[2]};
  double kst[3];
  oref_convolut_magic_t_per_self(n1i - 1, n2i - 1, n3i - 1, psir, y_c);
  oref_convolut_kinetic_per_T(n1i - 1, n2i - 1, n3i - 1, hh, x_c, y_c, kst);
  oref_analyse_per_self(n1, n2, n3, y_c, psir);
  oref_compress_and_accumulate_mixed(n1, n2, n3, nseg_c, keyg_c, keyv_c,
                                     nseg_f, keyg_f, keyv_f, psir, hpsi,
                                     hpsi_f);
  *ekin = kst[0] + kst[1] + kst[2];
  free(x_c);
  free(y_c);
  free(psir);
}

/* free branch: uncompress_forstandard + comb_grow_all (daubisf.f90:235-260)
 * -> apply_potential_lr_bounds (potential.f90:1049-1248) -> comb_shrink +
 * ConvolkineticT + compress_and_accumulate_standard
 * (locreg_operations.f90:2642-2675) */
static void locham_free(const int n[3], const int nfl[3], const int nfu[3],
                        const double hgrid[3], int nseg_c, int nvctr_c,
                        const int32_t *keyg_c, const int32_t *keyv_c,
                        int nseg_f, int nvctr_f, const int32_t *keyg_f,
                        const int32_t *keyv_f, const double *pot,
                        const double *psi, double *hpsi, double *ekin,
                        double *epot, const oref_confpot *cd) {
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const int n1i = 2 * n1 + 31, n2i = 2 * n2 + 31, n3i = 2 * n3 + 31;
  const size_t nc = (size_t)(n1 + 1) * (n2 + 1) * (n3 + 1);
  const size_t nf = (size_t)7 * (nfu[0] - nfl[0] + 1) *
                    (nfu[1] - nfl[1] + 1) * (nfu[2] - nfl[2] + 1);
  double *x_c = malloc(sizeof(double) * nc);
  double *y_c = malloc(sizeof(double) * nc);
  double *x_f = malloc(sizeof(double) * (nf ? nf : 1));
  double *y_f = malloc(sizeof(double) * (nf ? nf : 1));
  double *psir = malloc(sizeof(double) * (size_t)n1i * n2i * n3i);
  const double *psi_f = nvctr_f > 0 ? psi + nvctr_c : psi;
  double *hpsi_f = nvctr_f > 0 ? hpsi + nvctr_c : hpsi;
  oref_uncompress_forstandard(n1, n2, n3, nfl, nfu, nseg_c, keyg_c, keyv_c,
                              nseg_f, keyg_f, keyv_f, psi, psi_f, x_c, x_f);
  oref_comb_grow_all(n1, n2, n3, nfl, nfu, x_c, x_f, psir);
  const int ishift[3] = {0, 0, 0};
  oref_apply_potential_lr_conf(n1i, n2i, n3i, n1i, n2i, n3i, ishift, psir, pot,
                               cd, epot);
  oref_comb_shrink(n1, n2, n3, nfl, nfu, psir, y_c, y_f);
  oref_convolkinetic_T(n1, n2, n3, nfl, nfu, hgrid, x_c, x_f, y_c, y_f, ekin);
  oref_compress_and_accumulate_standard(n1, n2, n3, nfl, nfu, nseg_c, keyg_c,
                                        keyv_c, nseg_f, keyg_f, keyv_f, y_c,
                                        y_f, hpsi, hpsi_f);
  free(x_c);
  free(y_c);
  free(x_f);
  free(y_f);
  free(psir);
}

void oref_locham_per(const int n[3], const double hgrid[3], int nseg_c,
                     int nvctr_c, const int32_t *keyg_c, const int32_t *keyv_c,
                     int nseg_f, int nvctr_f, const int32_t *keyg_f,
                     const int32_t *keyv_f, const double *pot,
                     const double *psi, double *hpsi, double *ekin,
                     double *epot) {
  locham_per(n, hgrid, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f, nvctr_f, keyg_f,
             keyv_f, pot, psi, hpsi, ekin, epot, NULL);
}

void oref_locham_free(const int n[3], const int nfl[3], const int nfu[3],
                      const double hgrid[3], int nseg_c, int nvctr_c,
                      const int32_t *keyg_c, const int32_t *keyv_c,
                      int nseg_f, int nvctr_f, const int32_t *keyg_f,
                      const int32_t *keyv_f, const double *pot,
                      const double *psi, double *hpsi, double *ekin,
                      double *epot) {
  locham_free(n, nfl, nfu, hgrid, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f,
              nvctr_f, keyg_f, keyv_f, pot, psi, hpsi, ekin, epot, NULL);
}

/* the ket loop of LocalHamiltonianApplication, hpsiortho.f90:945-958, with
 * the orbital's confdata passed to psir_to_vpsi (orbital_basis.f90:830-831);
 * conf = NULL or potorder = 0: no confinement */
void oref_locham_many_conf(char geocode, const int n[3], const int nfl[3],
                           const int nfu[3], const double hgrid[3], int nseg_c,
                           int nvctr_c, const int32_t *keyg_c,
                           const int32_t *keyv_c, int nseg_f, int nvctr_f,
                           const int32_t *keyg_f, const int32_t *keyv_f,
                           const double *pot, int norb, const double *psi,
                           double *hpsi, double *ekin, double *epot,
                           const oref_confpot *conf) {
  const size_t ncoef = (size_t)nvctr_c + 7 * (size_t)nvctr_f;
  for (int iorb = 0; iorb < norb; ++iorb) {
    const oref_confpot *cd = conf ? conf + iorb : NULL;
    if (geocode == 'P')
      locham_per(n, hgrid, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f, nvctr_f,
                 keyg_f, keyv_f, pot, psi + ncoef * iorb, hpsi + ncoef * iorb,
                 ekin + iorb, epot + iorb, cd);
    else
      locham_free(n, nfl, nfu, hgrid, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f,
                  nvctr_f, keyg_f, keyv_f, pot, psi + ncoef * iorb,
                  hpsi + ncoef * iorb, ekin + iorb, epot + iorb, cd);
  }
}

void oref_locham_many(char geocode, const int n[3], const int nfl[3],
                      const int nfu[3], const double hgrid[3], int nseg_c,
                      int nvctr_c, const int32_t *keyg_c,
                      const int32_t *keyv_c, int nseg_f, int nvctr_f,
                      const int32_t *keyg_f, const int32_t *keyv_f,
                      const double *pot, int norb, const double *psi,
                      double *hpsi, double *ekin, double *epot) {
  oref_locham_many_conf(geocode, n, nfl, nfu, hgrid, nseg_c, nvctr_c, keyg_c,
                        keyv_c, nseg_f, nvctr_f, keyg_f, keyv_f, pot, norb, psi,
                        hpsi, ekin, epot, NULL);
}

/* daub_to_isf (orbitals/daubisf.f90:13-80): psir on the fine ISF grid of the
 * box (periodic: uncompress_per + convolut_magic_n_per, daubisf.f90:70-80;
 * free: uncompress_forstandard + comb_grow_all, daubisf.f90:38-55) */
void oref_daub_to_isf(char geocode, const int n[3], const int nfl[3],
                      const int nfu[3], int nseg_c, int nvctr_c,
                      const int32_t *keyg_c, const int32_t *keyv_c, int nseg_f,
                      int nvctr_f, const int32_t *keyg_f,
                      const int32_t *keyv_f, const double *psi, double *psir) {
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const double *psi_f = nvctr_f > 0 ? psi + nvctr_c : psi;
  if (geocode == 'P') {
    const size_t tot = (size_t)(2 * n1 + 2) * (2 * n2 + 2) * (2 * n3 + 2);
    double *x_c = malloc(sizeof(double) * tot);
    double *y_c = malloc(sizeof(double) * tot);
    oref_uncompress_per(n1, n2, n3, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f,
                        nvctr_f, keyg_f, keyv_f, psi, psi_f, x_c, psir);
    oref_convolut_magic_n_per(2 * n1 + 1, 2 * n2 + 1, 2 * n3 + 1, x_c, psir,
                              y_c);
    free(x_c);
    free(y_c);
  } else {
    const size_t nc = (size_t)(n1 + 1) * (n2 + 1) * (n3 + 1);
    const size_t nf = (size_t)7 * (nfu[0] - nfl[0] + 1) *
                      (nfu[1] - nfl[1] + 1) * (nfu[2] - nfl[2] + 1);
    double *x_c = malloc(sizeof(double) * nc);
    double *x_f = malloc(sizeof(double) * (nf ? nf : 1));
    oref_uncompress_forstandard(n1, n2, n3, nfl, nfu, nseg_c, keyg_c, keyv_c,
                                nseg_f, keyg_f, keyv_f, psi, psi_f, x_c, x_f);
    oref_comb_grow_all(n1, n2, n3, nfl, nfu, x_c, x_f, psir);
    free(x_c);
    free(x_f);
  }
}

/* local_partial_density (sumrho.f90:377-469) with partial_density
 * (sumrho.f90:472-560) for npsir = 1 and no slab scattering: for each
 * orbital, rho(:, spin) += hfac * psir^2, hfac = kwgt*occup/volume_element
 * given by the caller (sumrho.f90:425); rho is accumulated. */
void oref_density(char geocode, const int n[3], const int nfl[3],
                  const int nfu[3], int nseg_c, int nvctr_c,
                  const int32_t *keyg_c, const int32_t *keyv_c, int nseg_f,
                  int nvctr_f, const int32_t *keyg_f, const int32_t *keyv_f,
                  int norb, const double *psi, const double *hfac,
                  const int *spin, double *rho) {
  const int add = geocode == 'P' ? 2 : 31;
  const size_t tot = (size_t)(2 * n[0] + add) * (2 * n[1] + add) *
                     (2 * n[2] + add);
  const size_t ncoef = (size_t)nvctr_c + 7 * (size_t)nvctr_f;
  double *psir = malloc(sizeof(double) * tot);
  for (int iorb = 0; iorb < norb; ++iorb) {
    if (hfac[iorb] == 0.0) continue;
    oref_daub_to_isf(geocode, n, nfl, nfu, nseg_c, nvctr_c, keyg_c, keyv_c,
                     nseg_f, nvctr_f, keyg_f, keyv_f, psi + ncoef * iorb,
                     psir);
    double *r = rho + (spin ? (size_t)spin[iorb] * tot : 0);
    for (size_t i = 0; i < tot; ++i) r[i] += hfac[iorb] * (psir[i] * psir[i]);
  }
  free(psir);
}

/* ------------------------------------------------------------------------ */
/* Transfers between a localisation region and the global region (free BC),  */
/* locregs/locreg_operations.f90.  Local keys are in the local box (0:nl);   */
/* the local box starts at ns in the global box (0:ng).                      */
/* ------------------------------------------------------------------------ */

/* shift_locreg_indexes (locreg_operations.f90:2321-2380): local segment
 * bounds as 1-based linear indices of the global box */
static void shift_keys(const int nl[3], const int ng[3], const int ns[3],
                       int nseg, const int32_t *keyg_l, int32_t *mask) {
  const long long pl = (long long)(nl[0] + 1) * (nl[1] + 1);
  for (int s = 0; s < 2 * nseg; ++s) {
    long long t = keyg_l[s] - 1;
    const long long z = t / pl;
    t -= z * pl;
    const long long y = t / (nl[0] + 1);
    const long long x = t - y * (nl[0] + 1);
    mask[s] = (int32_t)((z + ns[2]) * (ng[0] + 1) * (long long)(ng[1] + 1) +
                        (y + ns[1]) * (ng[0] + 1) + (x + ns[0]) + 1);
  }
}

/* psi_to_locreg2 (locreg_operations.f90:2143-2318): lpsi = 0, then every
 * element of a local segment that overlaps a global segment is copied from
 * gpsi.  The local index advances with each copied element, as in the
 * reference (correct when every local point lies in the global region). */
void oref_psi_to_locreg2(const int nl[3], const int ng[3], const int ns[3],
                         int nseg_lc, int nseg_lf, int nvctr_lc, int nvctr_lf,
                         const int32_t *keyg_l, const int32_t *keyv_l,
                         int nseg_gc, int nseg_gf, int nvctr_gc,
                         const int32_t *keyg_g, const int32_t *keyv_g,
                         const double *gpsi, double *lpsi) {
  const int nseg = nseg_lc + nseg_lf;
  int32_t *mask = malloc(sizeof(int32_t) * 2 * (nseg ? nseg : 1));
  shift_keys(nl, ng, ns, nseg, keyg_l, mask);
  memset(lpsi, 0, sizeof(double) * ((size_t)nvctr_lc + 7 * (size_t)nvctr_lf));
  int isegstart = 0;
  for (int il = 0; il < nseg_lc; ++il) {
    const int lmin = mask[2 * il], lmax = mask[2 * il + 1];
    int istart = keyv_l[il] - 1;
    for (int ig = isegstart; ig < nseg_gc; ++ig) {
      const int gmin = keyg_g[2 * ig], gmax = keyg_g[2 * ig + 1];
      if (lmin > gmax) isegstart = ig;
      if (gmin > lmax) break;
      if (lmin > gmax || lmax < gmin) continue;
      const int offset = lmin - gmin > 0 ? lmin - gmin : 0;
      const int length = (lmax < gmax ? lmax : gmax) - (lmin > gmin ? lmin : gmin);
      for (int ix = 0; ix <= length; ++ix) {
        istart += 1;
        lpsi[istart - 1] = gpsi[keyv_g[ig] + offset + ix - 1];
      }
    }
  }
  const int start = nvctr_lc, gstart = nvctr_gc;
  isegstart = nseg_gc;
  for (int il = nseg_lc; il < nseg; ++il) {
    const int lmin = mask[2 * il], lmax = mask[2 * il + 1];
    int istart = keyv_l[il] - 1;
    for (int ig = isegstart; ig < nseg_gc + nseg_gf; ++ig) {
      const int gmin = keyg_g[2 * ig], gmax = keyg_g[2 * ig + 1];
      if (lmin > gmax) isegstart = ig;
      if (gmin > lmax) break;
      if (lmin > gmax || lmax < gmin) continue;
      const int offset = lmin - gmin > 0 ? lmin - gmin : 0;
      const int length = (lmax < gmax ? lmax : gmax) - (lmin > gmin ? lmin : gmin);
      for (int ix = 0; ix <= length; ++ix) {
        istart += 1;
        for (int igrid = 1; igrid <= 7; ++igrid)
          lpsi[start + (size_t)(istart - 1) * 7 + igrid - 1] =
              gpsi[gstart + (size_t)(keyv_g[ig] + offset + ix - 1) * 7 + igrid - 1];
      }
    }
  }
  free(mask);
}

/* Lpsi_to_global2 (locreg_operations.f90:1665-1837), nspin = 1: the
 * elements of the local segments (in global coordinates) are written into
 * psi where they overlap global segments; psi is otherwise left as it is.
 * Returns the number of local points placed (the reference stops when it
 * differs from nvctr_c + nvctr_f). */
int oref_lpsi_to_global2(const int nl[3], const int ng[3], const int ns[3],
                         int nseg_lc, int nseg_lf, int nvctr_lc, int nvctr_lf,
                         const int32_t *keyg_l, const int32_t *keyv_l,
                         int nseg_gc, int nseg_gf, int nvctr_gc,
                         const int32_t *keyg_g, const int32_t *keyv_g,
                         const double *lpsi, double *psi) {
  (void)nvctr_lf;
  const int nseg = nseg_lc + nseg_lf;
  int32_t *mask = malloc(sizeof(int32_t) * 2 * (nseg ? nseg : 1));
  shift_keys(nl, ng, ns, nseg, keyg_l, mask);
  int icheck = 0, isegstart = 0;
  for (int il = 0; il < nseg_lc; ++il) {
    const int lmin = mask[2 * il], lmax = mask[2 * il + 1];
    int istart = keyv_l[il] - 1;
    for (int ig = isegstart; ig < nseg_gc; ++ig) {
      const int gmin = keyg_g[2 * ig], gmax = keyg_g[2 * ig + 1];
      if (lmin > gmax) isegstart = ig;
      if (gmin > lmax) break;
      if (lmin > gmax) continue;
      const int offset = lmin - gmin > 0 ? lmin - gmin : 0;
      const int length = (lmax < gmax ? lmax : gmax) - (lmin > gmin ? lmin : gmin);
      icheck += length + 1;
      for (int ix = 0; ix <= length; ++ix) {
        istart += 1;
        psi[keyv_g[ig] + offset + ix - 1] = lpsi[istart - 1];
      }
    }
  }
  const int start = nvctr_lc, gstart = nvctr_gc;
  isegstart = nseg_gc;
  for (int il = nseg_lc; il < nseg; ++il) {
    const int lmin = mask[2 * il], lmax = mask[2 * il + 1];
    int istart = keyv_l[il] - 1;
    for (int ig = isegstart; ig < nseg_gc + nseg_gf; ++ig) {
      const int gmin = keyg_g[2 * ig], gmax = keyg_g[2 * ig + 1];
      if (lmin > gmax) isegstart = ig;
      if (gmin > lmax) break;
      if (lmin > gmax) continue;
      const int offset = lmin - gmin > 0 ? lmin - gmin : 0;
      const int length = (lmax < gmax ? lmax : gmax) - (lmin > gmin ? lmin : gmin);
      icheck += length + 1;
      for (int ix = 0; ix <= length; ++ix) {
        istart += 1;
        for (int igrid = 1; igrid <= 7; ++igrid)
          psi[gstart + (size_t)(keyv_g[ig] + offset + ix - 1) * 7 + igrid - 1] =
              lpsi[start + (size_t)(istart - 1) * 7 + igrid - 1];
      }
    }
  }
  free(mask);
  return icheck;
}

/* ------------------------------------------------------------------------ */
/* Descriptors, locregs/initlocregs.f90:1573-1686.                           */
/* logrid is (0:n1,0:n2,0:n3) column-major, nonzero = point present.         */
/* ------------------------------------------------------------------------ */

/* num_segkeys, initlocregs.f90:1573-1629 */
int oref_num_segkeys(int n1, int n2, int n3, int nl1, int nu1, int nl2,
                     int nu2, int nl3, int nu3, const uint8_t *logrid,
                     int *mseg, int *mvctr) {
  int nsrt = 0, nend = 0, mv = 0;
  (void)n3;
  for (int i3 = nl3; i3 <= nu3; ++i3)
    for (int i2 = nl2; i2 <= nu2; ++i2) {
      int pl = 0;
      for (int i1 = nl1; i1 <= nu1; ++i1) {
        const int lg =
            logrid[i1 + (size_t)(n1 + 1) * (i2 + (size_t)(n2 + 1) * i3)] != 0;
        if (lg) {
          ++mv;
          if (!pl) ++nsrt;
        } else if (pl) {
          ++nend;
        }
        pl = lg;
      }
      if (pl) ++nend;
    }
  if (nend != nsrt) return 1;
  *mseg = nend;
  *mvctr = mv;
  return 0;
}

/* segkeys, initlocregs.f90:1633-1686 */
int oref_segkeys(int n1, int n2, int n3, int nl1, int nu1, int nl2, int nu2,
                 int nl3, int nu3, const uint8_t *logrid, int mseg,
                 int32_t *keyg, int32_t *keyv) {
  int mvctr = 0, nsrt = 0, nend = 0;
  const int n1p1 = n1 + 1, np = n1p1 * (n2 + 1);
  (void)n3;
  for (int i3 = nl3; i3 <= nu3; ++i3)
    for (int i2 = nl2; i2 <= nu2; ++i2) {
      int pl = 0, ngridp = 0;
      for (int i1 = nl1; i1 <= nu1; ++i1) {
        ngridp = i3 * np + i2 * n1p1 + i1 + 1;
        const int lg =
            logrid[i1 + (size_t)(n1 + 1) * (i2 + (size_t)(n2 + 1) * i3)] != 0;
        if (lg) {
          ++mvctr;
          if (!pl) {
            if (nsrt >= mseg) return 2;
            keyg[2 * nsrt] = ngridp;
            keyv[nsrt] = mvctr;
            ++nsrt;
          }
        } else if (pl) {
          keyg[2 * nend + 1] = ngridp - 1;
          ++nend;
        }
        pl = lg;
      }
      if (pl) {
        keyg[2 * nend + 1] = ngridp;
        ++nend;
      }
    }
  return nend != nsrt ? 1 : 0;
}

/* ======================================================================
 * Convolution bounds, locregs/bounds.f90:555-1123.
 *
 * Restated with the reference's own mechanism: a flat logical work grid is
 * filled from one array's intervals (ib_to_logrid_*) and read back, with
 * its axes taken in a rotated order, as the first/last true index of every
 * row (ib_from_logrid_*).  The rotation is only a reinterpretation of the
 * flat grid's extents, as in the reference; the tests check every stage
 * against a direct per-axis formulation.
 * ====================================================================== */

/* Fortran ib(2, lo1:hi1, lo2:hi2) addressing */
#define IB(a, k, i, j, lo1, n1_, lo2) \
  (a)[(k) + 2 * (((i) - (lo1)) + (size_t)(n1_) * ((j) - (lo2)))]

int oref_bounds_layout(const int n[3], const int nfl[3], const int nfu[3],
                       int64_t off[18], int32_t dims[68]) {
  for (int d = 0; d < 3; ++d)
    if (nfl[d] > nfu[d]) return 1;
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const int l1 = nfl[0], u1 = nfu[0], l2 = nfl[1], u2 = nfu[1], l3 = nfl[2],
            u3 = nfu[2];
  /* bounds.f90:572-601 (allocation extents, same order as the types) */
  const int32_t d[17][4] = {
      {0, n2, 0, n3},                                  /* ibyz_c   */
      {0, n1, 0, n3},                                  /* ibxz_c   */
      {0, n1, 0, n2},                                  /* ibxy_c   */
      {0, n2, 0, n3},                                  /* ibyz_f   */
      {0, n1, 0, n3},                                  /* ibxz_f   */
      {0, n1, 0, n2},                                  /* ibxy_f   */
      {0, n3, -14, 2 * n1 + 16},                       /* ibzxx_c  */
      {-14, 2 * n1 + 16, -14, 2 * n2 + 16},            /* ibxxyy_c */
      {l2, u2, l3, u3},                                /* ibyz_ff  */
      {l3, u3, 2 * l1 - 14, 2 * u1 + 16},              /* ibzxx_f  */
      {2 * l1 - 14, 2 * u1 + 16, 2 * l2 - 14, 2 * u2 + 16}, /* ibxxyy_f */
      {-14, 2 * n3 + 16, 0, n1},                       /* ibzzx_c  */
      {-14, 2 * n2 + 16, -14, 2 * n3 + 16},            /* ibyyzz_c */
      {l1, u1, l2, u2},                                /* ibxy_ff  */
      {2 * l3 - 14, 2 * u3 + 16, l1, u1},              /* ibzzx_f  */
      {2 * l2 - 14, 2 * u2 + 16, 2 * l3 - 14, 2 * u3 + 16}, /* ibyyzz_f */
      {-14, 2 * n2 + 16, -14, 2 * n3 + 16}};           /* ibyyzz_r */
  int64_t o = 0;
  for (int k = 0; k < 17; ++k) {
    off[k] = o;
    for (int j = 0; j < 4; ++j) dims[4 * k + j] = d[k][j];
    o += 2 * (int64_t)(d[k][1] - d[k][0] + 1) * (d[k][3] - d[k][2] + 1);
  }
  off[17] = o;
  return 0;
}

/* make_bounds, bounds.f90:690-775 */
static void ob_make_bounds(int n1, int n2, int n3, const uint8_t *lg,
                           int32_t *ibyz, int32_t *ibxz, int32_t *ibxy) {
#define LG(i1, i2, i3) lg[(i1) + (size_t)(n1 + 1) * ((i2) + (size_t)(n2 + 1) * (i3))]
  for (int i3 = 0; i3 <= n3; ++i3)
    for (int i2 = 0; i2 <= n2; ++i2) {
      int a = 1000, b = -1000;
      for (int i1 = 0; i1 <= n1; ++i1)
        if (LG(i1, i2, i3)) { a = i1; break; }
      for (int i1 = n1; i1 >= 0; --i1)
        if (LG(i1, i2, i3)) { b = i1; break; }
      IB(ibyz, 0, i2, i3, 0, n2 + 1, 0) = a;
      IB(ibyz, 1, i2, i3, 0, n2 + 1, 0) = b;
    }
  for (int i3 = 0; i3 <= n3; ++i3)
    for (int i1 = 0; i1 <= n1; ++i1) {
      int a = 1000, b = -1000;
      for (int i2 = 0; i2 <= n2; ++i2)
        if (LG(i1, i2, i3)) { a = i2; break; }
      for (int i2 = n2; i2 >= 0; --i2)
        if (LG(i1, i2, i3)) { b = i2; break; }
      IB(ibxz, 0, i1, i3, 0, n1 + 1, 0) = a;
      IB(ibxz, 1, i1, i3, 0, n1 + 1, 0) = b;
    }
  for (int i2 = 0; i2 <= n2; ++i2)
    for (int i1 = 0; i1 <= n1; ++i1) {
      int a = 1000, b = -1000;
      for (int i3 = 0; i3 <= n3; ++i3)
        if (LG(i1, i2, i3)) { a = i3; break; }
      for (int i3 = n3; i3 >= 0; --i3)
        if (LG(i1, i2, i3)) { b = i3; break; }
      IB(ibxy, 0, i1, i2, 0, n1 + 1, 0) = a;
      IB(ibxy, 1, i1, i2, 0, n1 + 1, 0) = b;
    }
#undef LG
}

/* ib_to_logrid_inv, bounds.f90:895-915: logrid(-14+2nfl:2nfu+16, ndat) */
static void ob_to_logrid_inv(const int32_t *ib, uint8_t *lg, int nfl, int nfu,
                             long ndat) {
  const long w = 2L * (nfu - nfl) + 31;
  memset(lg, 0, (size_t)(w * ndat));
  for (long l = 0; l < ndat; ++l)
    for (int i = 2 * ib[2 * l] - 14; i <= 2 * ib[2 * l + 1] + 16; ++i)
      lg[(i - (2 * nfl - 14)) + w * l] = 1;
}

/* ib_from_logrid_inv, bounds.f90:919-950: logrid(ndat, ml1:mu1) */
static void ob_from_logrid_inv(int32_t *ib, const uint8_t *lg, int ml1,
                               int mu1, long ndat) {
  for (long i = 0; i < ndat; ++i) {
    ib[2 * i] = 1000;
    ib[2 * i + 1] = -1000;
    for (int i1 = ml1; i1 <= mu1; ++i1)
      if (lg[i + ndat * (i1 - ml1)]) { ib[2 * i] = i1; break; }
    for (int i1 = mu1; i1 >= ml1; --i1)
      if (lg[i + ndat * (i1 - ml1)]) { ib[2 * i + 1] = i1; break; }
  }
}

/* ib_to_logrid_rot, bounds.f90:1010-1028: logrid(ndat, -14+2nfl:2nfu+16) */
static void ob_to_logrid_rot(const int32_t *ib, uint8_t *lg, int nfl, int nfu,
                             long ndat) {
  const long w = 2L * (nfu - nfl) + 31;
  memset(lg, 0, (size_t)(w * ndat));
  for (long l = 0; l < ndat; ++l)
    for (int i = 2 * ib[2 * l] - 14; i <= 2 * ib[2 * l + 1] + 16; ++i)
      lg[l + ndat * (i - (2 * nfl - 14))] = 1;
}

/* ib_from_logrid, bounds.f90:1032-1063: logrid(ml1:mu1, ndat) */
static void ob_from_logrid(int32_t *ib, const uint8_t *lg, int ml1, int mu1,
                           long ndat) {
  const long w = mu1 - ml1 + 1;
  for (long i = 0; i < ndat; ++i) {
    ib[2 * i] = 1000;
    ib[2 * i + 1] = -1000;
    for (int i1 = ml1; i1 <= mu1; ++i1)
      if (lg[(i1 - ml1) + w * i]) { ib[2 * i] = i1; break; }
    for (int i1 = mu1; i1 >= ml1; --i1)
      if (lg[(i1 - ml1) + w * i]) { ib[2 * i + 1] = i1; break; }
  }
}

/* squares, bounds.f90:1097-1123: ib(2, 0:n2, 0:n3) made of 2x2 blocks */
static void ob_squares(int32_t *ib, int n2, int n3) {
  if (n2 == 0 || n3 == 0) return;
  for (int i3 = 0; i3 <= (n3 - 1) / 2; ++i3) {
    const int j3 = 2 * i3;
    for (int i2 = 0; i2 <= (n2 - 1) / 2; ++i2) {
      const int j2 = 2 * i2;
      int32_t *p[4] = {&IB(ib, 0, j2, j3, 0, n2 + 1, 0),
                       &IB(ib, 0, j2 + 1, j3, 0, n2 + 1, 0),
                       &IB(ib, 0, j2, j3 + 1, 0, n2 + 1, 0),
                       &IB(ib, 0, j2 + 1, j3 + 1, 0, n2 + 1, 0)};
      int mn = p[0][0], mx = p[0][1];
      for (int q = 1; q < 4; ++q) {
        if (p[q][0] < mn) mn = p[q][0];
        if (p[q][1] > mx) mx = p[q][1];
      }
      for (int q = 0; q < 4; ++q) {
        p[q][0] = mn;
        p[q][1] = mx;
      }
    }
  }
}

/* squares_1d, bounds.f90:1069-1091: ib(2, nfl2:nfu2, nfl3:nfu3) made of
 * pairs along the first axis (C division truncates like Fortran's) */
static void ob_squares_1d(int32_t *ib, int nfl2, int nfu2, int nfl3, int nfu3) {
  const int w = nfu2 - nfl2 + 1;
  for (int i3 = nfl3; i3 <= nfu3; ++i3)
    for (int i2 = nfl2 / 2; i2 <= (nfu2 - 1) / 2; ++i2) {
      const int j2 = 2 * i2;
      int32_t *a = &IB(ib, 0, j2, i3, nfl2, w, nfl3);
      int32_t *b = &IB(ib, 0, j2 + 1, i3, nfl2, w, nfl3);
      const int mn = a[0] < b[0] ? a[0] : b[0];
      const int mx = a[1] > b[1] ? a[1] : b[1];
      a[0] = b[0] = mn;
      a[1] = b[1] = mx;
    }
}

/* make_ib_inv, bounds.f90:868-891 */
static void ob_make_ib_inv(uint8_t *lg, const int32_t *ibxy, int32_t *ibzzx,
                           int32_t *ibyyzz, int l1, int u1, int l2, int u2,
                           int l3, int u3) {
  long nt = (long)(u1 - l1 + 1) * (u2 - l2 + 1);
  ob_to_logrid_inv(ibxy, lg, l3, u3, nt);
  nt = (2L * (u3 - l3) + 31) * (u1 - l1 + 1);
  ob_from_logrid_inv(ibzzx, lg, l2, u2, nt);
  ob_to_logrid_inv(ibzzx, lg, l2, u2, nt);
  nt = (2L * (u2 - l2) + 31) * (2L * (u3 - l3) + 31);
  ob_from_logrid_inv(ibyyzz, lg, l1, u1, nt);
}

/* make_ib_c, bounds.f90:954-980 (squares ibyz in place, as the reference) */
static void ob_make_ib_c(uint8_t *lg, int32_t *ibyz, int32_t *ibzxx,
                         int32_t *ibxxyy, int n1, int n2, int n3) {
  ob_squares(ibyz, n2, n3);
  long nt = (long)(n2 + 1) * (n3 + 1);
  ob_to_logrid_rot(ibyz, lg, 0, n1, nt);
  nt = (long)(n3 + 1) * (2 * n1 + 31);
  ob_from_logrid(ibzxx, lg, 0, n2, nt);
  ob_squares(ibzxx, n3, 2 * n1 + 30);
  ob_to_logrid_rot(ibzxx, lg, 0, n2, nt);
  nt = (long)(2 * n1 + 31) * (2 * n2 + 31);
  ob_from_logrid(ibxxyy, lg, 0, n3, nt);
  ob_squares(ibxxyy, 2 * n1 + 30, 2 * n2 + 30);
}

/* make_ib, bounds.f90:984-1006 */
static void ob_make_ib(uint8_t *lg, const int32_t *ibyz, int32_t *ibzxx,
                       int32_t *ibxxyy, int l1, int u1, int l2, int u2, int l3,
                       int u3) {
  long nt = (long)(u2 - l2 + 1) * (u3 - l3 + 1);
  ob_to_logrid_rot(ibyz, lg, l1, u1, nt);
  nt = (long)(u3 - l3 + 1) * (2 * (u1 - l1) + 31);
  ob_from_logrid(ibzxx, lg, l2, u2, nt);
  ob_to_logrid_rot(ibzxx, lg, l2, u2, nt);
  nt = (long)(2 * (u1 - l1) + 31) * (2 * (u2 - l2) + 31);
  ob_from_logrid(ibxxyy, lg, l3, u3, nt);
}

/* wfd_to_logrids (bounds.f90:616-688), make_bounds, make_all_ib
 * (bounds.f90:777-865) as called by locreg_bounds (bounds.f90:555-613) */
int oref_locreg_bounds(const int n[3], const int nfl[3], const int nfu[3],
                       int nseg_c, const int32_t *keyg_c, int nseg_f,
                       const int32_t *keyg_f, int32_t *out) {
  int64_t off[18];
  int32_t dims[68];
  if (oref_bounds_layout(n, nfl, nfu, off, dims)) return 1;
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const size_t M = (size_t)(n1 + 1) * (n2 + 1) * (n3 + 1);
  uint8_t *lgc = calloc(M, 1), *lgf = calloc(M, 1);
  for (int f = 0; f < 2; ++f) {
    const int nseg = f ? nseg_f : nseg_c;
    const int32_t *kg = f ? keyg_f : keyg_c;
    uint8_t *lg = f ? lgf : lgc;
    for (int s = 0; s < nseg; ++s) {
      const long j0 = kg[2 * s] - 1, len = kg[2 * s + 1] - kg[2 * s];
      for (long i = 0; i <= len; ++i) lg[j0 + i] = 1;
    }
  }
  int32_t *A[17];
  for (int k = 0; k < 17; ++k) A[k] = out + off[k];
  ob_make_bounds(n1, n2, n3, lgc, A[0], A[1], A[2]);
  ob_make_bounds(n1, n2, n3, lgf, A[3], A[4], A[5]);
  free(lgc);
  free(lgf);
  const int l1 = nfl[0], u1 = nfu[0], l2 = nfl[1], u2 = nfu[1], l3 = nfl[2],
            u3 = nfu[2];
  uint8_t *big = malloc((size_t)(2 * n1 + 31) * (2 * n2 + 31) * (2 * n3 + 31));
  /* shrink: ibxy_ff = ibxy_f on the fine box */
  for (int i2 = l2; i2 <= u2; ++i2)
    for (int i1 = l1; i1 <= u1; ++i1)
      for (int k = 0; k < 2; ++k)
        IB(A[13], k, i1, i2, l1, u1 - l1 + 1, l2) = IB(A[5], k, i1, i2, 0, n1 + 1, 0);
  ob_make_ib_inv(big, A[2], A[11], A[12], 0, n1, 0, n2, 0, n3);
  ob_make_ib_inv(big, A[13], A[14], A[15], l1, u1, l2, u2, l3, u3);
  /* real space */
  const int w2 = 2 * n2 + 31;
  for (int i3 = -14; i3 <= 2 * n3 + 16; ++i3)
    for (int i2 = -14; i2 <= 2 * n2 + 16; ++i2) {
      const int a = IB(A[12], 0, i2, i3, -14, w2, -14);
      if (a != 1000) {
        IB(A[16], 0, i2, i3, -14, w2, -14) = 2 * a;
        IB(A[16], 1, i2, i3, -14, w2, -14) = 2 * IB(A[12], 1, i2, i3, -14, w2, -14) + 30;
      } else {
        IB(A[16], 0, i2, i3, -14, w2, -14) = 1000;
        IB(A[16], 1, i2, i3, -14, w2, -14) = -1000;
      }
    }
  ob_squares(A[16], 2 * n2 + 30, 2 * n3 + 30);
  /* grow: ibyz_ff = ibyz_f on the fine box */
  for (int i2 = l2; i2 <= u2; ++i2)
    for (int i3 = l3; i3 <= u3; ++i3)
      for (int k = 0; k < 2; ++k)
        IB(A[8], k, i2, i3, l2, u2 - l2 + 1, l3) = IB(A[3], k, i2, i3, 0, n2 + 1, 0);
  ob_make_ib_c(big, A[0], A[6], A[7], n1, n2, n3);
  ob_make_ib(big, A[8], A[9], A[10], l1, u1, l2, u2, l3, u3);
  ob_squares_1d(A[10], 2 * l1 - 14, 2 * u1 + 16, 2 * l2 - 14, 2 * u2 + 16);
  free(big);
  return 0;
}
#undef IB

/* ======================================================================
 * Preconditioner: the GPU part of preconditionall_OCL
 * (bigdft/src/interface_ocl.f90:622-880), ocl_preconditioner_generic_k
 * (bigdft/src/OpenCL/Preconditioner.c:245-358), real wavefunctions.
 *
 * apply_hp (Preconditioner.c:198-243): y = S Ana (K + c) Syn S x on the
 * compressed vector, S the diagonal scaling of each of the 8 component
 * types (scale_psi/uncompress_scale/compress_scale, OpenCL/Uncompress.c:
 * 377-470: scal_t = 1/sqrt(sum_d (bit_d(t) ? B2 : A2) / (8 h_d^2) + c),
 * h = hgrid/2), Syn/Ana the 3D wavelet synthesis/analysis (periodic, or
 * grown by 7 coarse points per side and shrunk back for free BC:
 * syn_self_d_generic / ana_self_d_generic, OpenCL/Wavelet.c:166-314), K
 * the fine-grid kinetic (kinetic_d_generic, OpenCL/Kinetic.c:240) with
 * spacing h and zero outside the grid for free BC.
 * ====================================================================== */
#define PREC_A2 3.55369228991319019
#define PREC_B2 24.8758460293923314

void oref_precond_scal(const double hgrid[3], double c, double scal[8]) {
  const double h1 = 0.5 * hgrid[0], h2 = 0.5 * hgrid[1], h3 = 0.5 * hgrid[2];
  const double hh[3] = {0.125 / (h1 * h1), 0.125 / (h2 * h2), 0.125 / (h3 * h3)};
  for (int t = 0; t < 8; ++t) {
    double s = c;
    for (int d = 0; d < 3; ++d) s += ((t >> d) & 1 ? PREC_B2 : PREC_A2) * hh[d];
    scal[t] = 1.0 / sqrt(s);
  }
}

/* free-BC synthesis of one axis with the full support: x(0:2n+1, ndat),
 * c in 0:n, d in n+1:2n+1 -> y(ndat, 0:2n+14-1), fine index j-7 at row j:
 * y(j) = sum_k ch(j-2k) c(k) + cg(j-2k) d(k), j-2k in -7..8 */
static void syngrow_rot(int n, long ndat, const double *x, double *y) {
  const int np1 = n + 1, m = 2 * n + 16;
#pragma omp parallel for schedule(static)
  for (long jd = 0; jd < ndat; ++jd) {
    const double *xj = x + (size_t)(2 * np1) * jd;
    for (int r = 0; r < m; ++r) {
      const int j = r - 7;
      double s = 0.0;
      for (int k = 0; k < np1; ++k) {
        const int l = j - 2 * k;
        if (l >= -7 && l <= 8) s += CHf(l) * xj[k] + CGf(l) * xj[np1 + k];
      }
      y[jd + (size_t)ndat * r] = s;
    }
  }
}

/* its adjoint, the analysis restricted to the box: y(0:2n+15, ndat) (row j
 * = fine index j-7) -> x(ndat, 0:2n+1): c(i) = sum_l ch(l) y(2i+l) */
static void anashrink_rot(int n, long ndat, const double *y, double *x) {
  const int np1 = n + 1, m = 2 * n + 16;
#pragma omp parallel for schedule(static)
  for (long jd = 0; jd < ndat; ++jd) {
    const double *yj = y + (size_t)m * jd;
    for (int i = 0; i < np1; ++i) {
      double ci = 0.0, di = 0.0;
      for (int l = -7; l <= 8; ++l) {
        const double v = yj[2 * i + l + 7];
        ci += CHf(l) * v;
        di += CGf(l) * v;
      }
      x[jd + (size_t)ndat * i] = ci;
      x[jd + (size_t)ndat * (np1 + i)] = di;
    }
  }
}

/* y += K x on an m1 x m2 x m3 fine grid, periodic or zero outside */
static void kinetic_fine(int m1, int m2, int m3, int per, const double h[3],
                         const double *x, double *y) {
  double fil[3][29];
  for (int d = 0; d < 3; ++d) {
    const double scale = -0.5 / (h[d] * h[d]);
    for (int l = 0; l <= 14; ++l) fil[d][14 + l] = fil[d][14 - l] = KIN_A[l] * scale;
  }
  const int m[3] = {m1, m2, m3};
  const size_t st[3] = {1, (size_t)m1, (size_t)m1 * m2};
#pragma omp parallel for schedule(static)
  for (int i3 = 0; i3 < m3; ++i3)
    for (int i2 = 0; i2 < m2; ++i2)
      for (int i1 = 0; i1 < m1; ++i1) {
        const int ii[3] = {i1, i2, i3};
        const size_t p = i1 + st[1] * i2 + st[2] * i3;
        double tt = 0.0;
        for (int d = 0; d < 3; ++d)
          for (int l = -14; l <= 14; ++l) {
            int q = ii[d] + l;
            if (per) q = pmod(q, m[d]);
            else if (q < 0 || q >= m[d]) continue;
            tt += fil[d][l + 14] * x[p + st[d] * (size_t)(q - ii[d])];
          }
        y[p] += tt;
      }
}

/* apply_hp_generic_k (Preconditioner.c:198-243), nspinor = 1 */
void oref_precond_apply(char geocode, const int n[3], const double hgrid[3],
                        double c, int nseg_c, int nvctr_c, const int32_t *keyg_c,
                        const int32_t *keyv_c, int nseg_f, int nvctr_f,
                        const int32_t *keyg_f, const int32_t *keyv_f,
                        const double *x, double *y) {
  const int n1 = n[0], n2 = n[1], n3 = n[2], per = geocode == 'P';
  double scal[8];
  oref_precond_scal(hgrid, c, scal);
  const size_t ncoef = (size_t)nvctr_c + 7 * (size_t)nvctr_f;
  double *xs = malloc(sizeof(double) * (ncoef ? ncoef : 1));
  for (size_t k = 0; k < (size_t)nvctr_c; ++k) xs[k] = scal[0] * x[k];
  for (size_t k = 0; k < (size_t)nvctr_f; ++k)
    for (int t = 0; t < 7; ++t)
      xs[nvctr_c + 7 * k + t] = scal[t + 1] * x[nvctr_c + 7 * k + t];
  const int e = per ? 2 : 16;  /* fine points per axis: 2(n+1) or 2(n+1)+14 */
  const int m1 = 2 * n1 + e, m2 = 2 * n2 + e, m3 = 2 * n3 + e;
  const size_t nw = (size_t)8 * (n1 + 1) * (n2 + 1) * (n3 + 1);
  const size_t ng = (size_t)m1 * m2 * m3;
  double *psig = malloc(sizeof(double) * nw);
  double *a = malloc(sizeof(double) * ng), *b = malloc(sizeof(double) * ng);
  oref_uncompress_psig(n1, n2, n3, nseg_c, keyg_c, keyv_c, nseg_f, keyg_f, keyv_f, xs,
                       xs + nvctr_c, psig);
  if (per) {
    memcpy(a, psig, sizeof(double) * nw);
    oref_synthese_per_self(n1, n2, n3, a, b);            /* result in b */
  } else {
    syngrow_rot(n1, (long)(2 * n2 + 2) * (2 * n3 + 2), psig, a);
    syngrow_rot(n2, (long)(2 * n3 + 2) * m1, a, b);
    syngrow_rot(n3, (long)m1 * m2, b, a);
    memcpy(b, a, sizeof(double) * ng);
  }
  const double hf[3] = {0.5 * hgrid[0], 0.5 * hgrid[1], 0.5 * hgrid[2]};
  for (size_t i = 0; i < ng; ++i) a[i] = c * b[i];
  kinetic_fine(m1, m2, m3, per, hf, b, a);               /* a = (c + K) b */
  if (per) {
    oref_analyse_per_self(n1, n2, n3, a, psig);           /* result in psig */
  } else {
    anashrink_rot(n1, (long)m2 * m3, a, b);
    anashrink_rot(n2, (long)m3 * (2 * n1 + 2), b, a);
    anashrink_rot(n3, (long)(2 * n1 + 2) * (2 * n2 + 2), a, psig);
  }
  memset(y, 0, sizeof(double) * ncoef);
  oref_compress_and_accumulate_mixed(n1, n2, n3, nseg_c, keyg_c, keyv_c, nseg_f, keyg_f,
                                     keyv_f, psig, y, y + nvctr_c);
  for (size_t k = 0; k < (size_t)nvctr_c; ++k) y[k] *= scal[0];
  for (size_t k = 0; k < (size_t)nvctr_f; ++k)
    for (int t = 0; t < 7; ++t) y[nvctr_c + 7 * k + t] *= scal[t + 1];
  free(xs);
  free(psig);
  free(a);
  free(b);
}

static double vdot(size_t n, const double *a, const double *b) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* ocl_preconditioner_generic_k (Preconditioner.c:245-358), nspinor = 1:
 * x (the guess from precondition_preconditioner) is overwritten by
 * S x_ncong; b is the scaled right-hand side (overwritten as work, as the
 * reference does with psi_c_b) */
void oref_precond_cg(char geocode, const int n[3], const double hgrid[3], double c,
                     int ncong, int nseg_c, int nvctr_c, const int32_t *keyg_c,
                     const int32_t *keyv_c, int nseg_f, int nvctr_f,
                     const int32_t *keyg_f, const int32_t *keyv_f, double *x,
                     double *b) {
  const size_t nc = (size_t)nvctr_c + 7 * (size_t)nvctr_f;
  double *r = malloc(sizeof(double) * (nc ? nc : 1));
  double *d = malloc(sizeof(double) * (nc ? nc : 1));
#define APPLY(in, out)                                                             \
  oref_precond_apply(geocode, n, hgrid, c, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f, \
                     nvctr_f, keyg_f, keyv_f, in, out)
  APPLY(x, d);
  for (size_t i = 0; i < nc; ++i) r[i] = b[i] - d[i];
  memcpy(d, r, sizeof(double) * nc);
  double ns = vdot(nc, r, r);
  for (int it = 0; it < ncong; ++it) {
    APPLY(d, b);
    const double alpha = ns / vdot(nc, d, b);
    for (size_t i = 0; i < nc; ++i) x[i] += alpha * d[i];
    if (it != ncong - 1) {
      for (size_t i = 0; i < nc; ++i) r[i] -= alpha * b[i];
      const double ns2 = vdot(nc, r, r);
      const double beta = ns2 / ns;
      for (size_t i = 0; i < nc; ++i) d[i] = beta * d[i] + r[i];
      ns = ns2;
    }
  }
#undef APPLY
  double scal[8];
  oref_precond_scal(hgrid, c, scal);
  for (size_t k = 0; k < (size_t)nvctr_c; ++k) x[k] *= scal[0];
  for (size_t k = 0; k < (size_t)nvctr_f; ++k)
    for (int t = 0; t < 7; ++t) x[nvctr_c + 7 * k + t] *= scal[t + 1];
  free(r);
  free(d);
}
