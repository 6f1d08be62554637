/*
 * CPU oracle for the RPA response path (FP64 restatement of the reference
 * Fortran loops).  TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it as the checker.  The product library
 * (librpa_hip.so) never links, loads or calls anything in this directory.
 *
 * Conventions (shared with include/rpa_hip.h):
 *   - arrays use Fortran column-major layout with the reference's shapes,
 *     e.g. E(NkI,Nband) is E[ik + NkI*n];
 *   - index maps that cross this interface are 1-based (K1, K2, iG0, lf_idx);
 *   - complex numbers are interleaved (re, im) doubles.
 */
#ifndef RPA_ORACLE_H
#define RPA_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex orc_z;

/* Vertex index-map semantics.
 *   ORC_MAP_DIRECT : drivers (Sloss.f90:907-946, Optabs.f90:1399-1449): the
 *                    first NGd coefficients, coefficient index == global G index.
 *   ORC_MAP_IGK    : Modules/vertices.f90:88-189 (_v2): the whole k-sphere,
 *                    coefficient i of k has G = G(igk_k(i,k)).               */
enum { ORC_MAP_DIRECT = 0, ORC_MAP_IGK = 1 };

typedef struct orc_problem {
  int nki;            /* IBZ k-points (== Ntot, Nsym = 1)                     */
  int nband;
  int npwx;           /* stride of the coefficient arrays                      */
  const int *ngk;     /* [nki] plane waves of each k                          */
  const int *mill_k;  /* (3, npwx, nki) Miller indices of each coefficient     */
  const double *evc;  /* (2, npwx, nband, nki) coefficients, complex128        */
  int ngm;            /* density G vectors                                     */
  const int *mill_g;  /* (3, ngm)                                              */
  const double *kc;   /* KC(3,3): cartesian = KC * crystal                     */
  const double *E;    /* E(nki, nband), Hartree, scissor already applied       */
  int ntot;           /* FBZ k-points                                          */
  const double *ktot; /* (3, ntot), 2pi/a0 units                               */
  const int *K1;      /* [ntot] 1-based IBZ index of k                         */
  const int *K2;      /* [ntot] 1-based IBZ index of k+q                       */
  const int *iG0;     /* [ntot] 1-based global G index of G0                   */
  double q[3];        /* 2pi/a0 units                                          */
  int nlf;
  const int *lf_idx;  /* [nlf] 1-based global G index of each local field      */
  int map_mode;       /* ORC_MAP_*                                             */
  int ngd;            /* NGd (direct mode)                                     */
  double eps;         /* G-vector match tolerance                              */
  /* point-group unfolding (Nsym > 1); R1 == NULL means the identity for
   * every k.  R, RI: (48,3,3) cartesian operations and their inverses as
   * PointR gives them; R1, R2: [ntot] 1-based operation of k and of k+q.    */
  int nsym;
  const double *R, *RI;
  const int *R1, *R2;
} orc_problem;

/* ---- host setup restatements (a3-a5) ---------------------------------- */
int  orc_gen_glf(const char *lf, double ecut, double gcar, int ng, const double *g,
                 const int *mill_g, int nlfd, int *nlf, double *glf, int *parg,
                 int *idx);
void orc_occupation(double ei, double ef, double T, double *expo, double *f);
int  orc_gen_fbz(int nki, int nsym, double eps, const double *ki, const double *R,
                 int *ntot, double *ktot);
void orc_find_min_q(int ntot, const double *ktot, int iq, double q[3]);
int  orc_find_kq(const double kq[3], double eps, int nki, int ntot, int ng,
                 const double *ktot, const double *ki, const double *g,
                 int *iG0, int *K2);
void orc_gcart(int n, const int *mill, const double *kc, double *g);
/* PointR's inverses (Chi_RPA/PointR.f90:119-138): RI(i) = gjel(R(i)). */
void orc_point_inverse(int nsym, const double *R, double *RI);
/* findKinIBZ (Optabs.f90:776-819): ik 1-based FBZ index; R1, K1 1-based.
 * Returns 1 when k is not in the IBZ (the reference stops).               */
int  orc_find_kin_ibz(int ik, int nki, int nsym, double eps, const double k[3],
                      const double *RI, const double *ki, int *R1, int *K1);
/* findKQinBZ with symmetry (Optabs.f90:822-869): iG0, R2, K2 1-based;
 * returns 1 (not in the FBZ) or 2 (not in the IBZ) as the reference.       */
int  orc_find_kq_sym(const double kq[3], double eps, int nsym, int nki, int ntot, int ng,
                     const double *ktot, const double *ki, const double *g, const double *RI,
                     int *iG0, int *R2, int *K2);
int  orc_gk_sort(const double k[3], int ngm, const double *g, double gcutw,
                 int npwx, int *ngk, int *igk);

/* ---- a6: vertex index map --------------------------------------------- */
void orc_gfast_linear(int n1, const double *l1, int n2, const double *l2, int nlf,
                      const double *glf, const double g0[3], double eps, int *gfast);
int  orc_gfast_miller(int n1, const int *m1, int n2, const int *m2, int nlf,
                      const int *mlf, const int m0[3], int *gfast);

/* ---- a7/a8: vertices for every transition ------------------------------ */
/* kind: 0 charge (Sloss), 1..3 current along x/y/z (Optabs), 4 = yz pair.
 * Transition order: k (FBZ), then n, then m over the band windows
 * [n_lo,n_hi) x [m_lo,m_hi).  rho1/rho2 are (nlf, ntrans) complex.          */
int  orc_vertices(const orc_problem *p, int kind, double gcar, int n_lo, int n_hi,
                  int m_lo, int m_hi, orc_z *rho1, orc_z *rho2);

/* ---- a9/a10: spectral-function accumulation ----------------------------- */
typedef struct orc_sloss_params {
  int nocc, no;
  double domega, eta, vcell;
} orc_sloss_params;
int  orc_sloss_s0(const orc_problem *p, const orc_sloss_params *sp, double *S0,
                  int nthreads);

typedef struct orc_optabs_params {
  int nval, no, pol;   /* pol: 1 x, 2 y, 3 z, 4 yz */
  double domega, eta_inter, lor_cut, df_cut, efermi, T, vcell, gcar;
} orc_optabs_params;
int  orc_optabs_s0(const orc_problem *p, const orc_optabs_params *op, orc_z *S0,
                   orc_z *Qeff, int nthreads);

/* ---- a12/a14/a20: Kramers-Kronig stencils -------------------------------- */
void orc_kk_optabs(int no, int nlf, double domega, const orc_z *S0, orc_z *Pi0);
void orc_kk_sloss(int no, int nlf, double domega, const double *S0, orc_z *chi0);
void orc_w1w2(int io, int no, int nlf, int iG, int jG, double domega,
              const double *S0, double *W1, double *W2);
void orc_kk_module_real(int io, int no, int nlf, int iG, int jG, double domega,
                        const double *S0m, double *re, double *im);

/* ---- a15-a19: Coulomb, Dyson, W ---------------------------------------- */
void orc_genV(double eps, double qx, double qy, double c0, int nlf, const int *parg,
              const double *glf, const double *glfv, double *V);
void orc_gjel(orc_z *a, int n, int np, orc_z *b, int m, int mp);
void orc_dyson_w(int nlf, const orc_z *chi0, const double *V, orc_z *epsinv,
                 orc_z *chi, orc_z *W);

/* ---- a13: Pi assembly (Optabs calc=2) ----------------------------------- */
void orc_pi_assembly(int no, int nlf, double domega, double eta_intra, double c0,
                     const orc_z *Pi0, const orc_z *Qeff, orc_z *pi11,
                     double *sig_inter, double *sig_intra);

/* ---- f4: Fock kernel of Chi_ladd, Chi_ladd_up/BSE.f90:305, 542, 804-816 ------- */
void orc_fock_kernel(int nk, int nlf, int nq, int ntot, double gcar, double qmin, double c0,
                     double vcell, const double *kbz, const orc_z *wt, const orc_z *mnm,
                     orc_z *fockk);

/* ---- f4: ladder (BSE) frequency loop, Chi_ladd_up/BSE.f90:1116-1176 ---------- */
/* fockk(nk,nk), current(2,nk) complex, delta(nk) -> chi_ladd(2,no),
 * chi_para(2,no) complex; FP64 throughout, fock inverted by orc_gjel.      */
void orc_ladder(int nk, int no, double omin, double domega, double eta, int ntot, double vcell,
                double valley, const orc_z *fockk, const double *delta, const orc_z *current,
                orc_z *chi_ladd, orc_z *chi_para);

/* ---- f4: GW propagators ------------------------------------------------------ */
/* genGammaPropagator (Modules/propagators.f90:11-108) for one series D(No):
 * gp, gm (No-1) complex, *kks, *skk the Kramers-Kronig check sums.        */
void orc_gamma_propagator(int no, double domega, const orc_z *D, orc_z *gp, orc_z *gm,
                          double *kks, double *skk);

/* ---- f3: Photon.x per (line, frequency) point ----------------------------- */
/* LU with partial pivoting and the inverse from it (zgetrf + zgetri, the
 * reference's invert, BSE/Photonew.f90:388-462); a is n x n column-major,
 * overwritten by its inverse.  Returns 0, or k > 0 for a zero pivot at k.  */
int  orc_lu_invert(orc_z *a, int n);
/* genD0 (Photonew.f90:618-663): D(N,N,4) = Dxx0, Dyy0, Dzz0, Dyz0 (= Dzy0). */
void orc_photon_d0(double q, double o, double eta, double c0, int nlf, const int *parg,
                   const double *glfz, orc_z *D);
/* One point of the angle / q / frequency loop (Photonew.f90:229-330):
 * pi0(N,N,5) = Pixx0, Piyy0, Piyz0, Pizy0, Pizz0 at this frequency;
 * pi(N,N,5) the screened blocks, sums(5) the sumIntScreenedPi values
 * (xx, yy, zz, yz, zy), spec(2) = A_p, R_p of genSpectra.  Returns the
 * orc_lu_invert status of TS, then of TP.                                 */
int  orc_photon_point(double q, double o, double eta, double c0, double theta, int nlf,
                      const int *parg, const double *glfz, const orc_z *pi0, orc_z *pi,
                      orc_z *sums, double *spec);

/* ---- f2: interpolate.x dense-grid response ------------------------------ */
void orc_interpolate(int no, int nlf, int no_interp, double domega_tot, double eta_intra,
                     const orc_z *Pi0, const orc_z *Qeff, orc_z *pi_interp, orc_z *pi_intra);

#ifdef __cplusplus
}
#endif
#endif
