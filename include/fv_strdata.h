/*
 * fv_strdata.h — host data model of the octree finite-volume solver.
 *
 * The reference's host structs (strdata.h:60-540) restated with the same type names,
 * field names, field order and field types, so that their layout is the reference's
 * byte for byte (x86-64, the MPI ABI below): a RUN built by the reference's host code
 * (its forest, leaves and scratch) can be handed to RHS(run), RK(irk, run),
 * faceinterp(run) and gradcalc(run) of this library, and the library's own forests
 * build RUNs of the same shape.  tests/test_abi_layout.py checks every field's offset
 * and size against the reference's declarations.
 *
 * MPI: CON begins with an MPI_Comm.  It is FV_MPI_COMM here, an int by default, which
 * is MPI_Comm in MPICH and its derivatives (the mpi.h of this image); define
 * FV_MPI_COMM before including this header to match another MPI (OpenMPI: a pointer).
 * The hot path has no MPI: the library never reads it.
 *
 * struct INFO is declared and never defined by the reference, and no reference file
 * reads or writes RUN::info.  This library defines it as the per-RUN state of the GPU
 * binding (host/bind.c): fv_gpu_attach() stores it in run->info.
 *
 * Ownership: every struct reachable from a RUN built by fvh_forest_create() is owned by
 * that forest handle and freed by fvh_forest_destroy().
 */
#ifndef FV_STRDATA_H
#define FV_STRDATA_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef FV_MPI_COMM
#define FV_MPI_COMM int
#endif

struct INFO;

/* node coordinates (strdata.h:62-67) */
typedef struct ND {
  int num;
  double x, y, z;
} ND;

/* face descriptor (strdata.h:69-74); bc: 0 interior, 1 inlet, 2 outflow (transmissive),
 * 3 reflective, 4 no-slip wall (3 and 4 call functions the reference never defines,
 * bc.c:38,42) */
typedef struct FC {
  int num;
  int type;
  int nnds;
  int bc;
} FC;

/* per-leaf state (strdata.h:76-93): S, SN, RHS hold neq doubles, SF[6][neq] the MUSCL
 * face states, SG[neq+1][3] the Green-Gauss gradients (conserved variables, then p) */
typedef struct LEAF {
  int gnum;
  double *RHS;
  double *S;
  double *SN;
  double **SF;
  double **SG;
  double crith;
  double **flux_face;
  int *flux_flag;
  int inner;
} LEAF;

/* a root of the partitioned mesh (strdata.h:96-102) */
typedef struct TREE {
  int type;
  int *vrtx;
  int *conf;
  int part;
  struct BRANCH *brch;
} TREE;

/* tree node, branch or leaf (strdata.h:104-148) */
typedef struct BRANCH {
  int igroup;
  int part;
  int oldpart;
  int level;             /* roots are level 1; the node outside the domain (glob) level 0 */
  int type;              /* 0 = hexahedron */
  int nkeens;
  int nkids;
  int nlnd;
  int nlfc;
  int tot_neigs;
  int *adrs;
  int split;
  int merge;
  int tag;
  int hangn;
  struct BRANCH *prnt;
  int root;
  struct BRANCH *next;
  struct BRANCH *lnxt;   /* leaf list, depth first = Morton order within a root */
  struct BRANCH *lprv;
  struct BRANCH ***neigtr;  /* neigtr[f][k]: k-th neighbour across face f */
  int **neignd;
  int *neigfc;              /* the neighbour's face id facing this leaf */
  int *neigag;
  int *lfc_neigs;           /* 1 same-or-coarser (or domain side), 4 finer */
  int *nsfc;                /* max(1, lfc_neigs) */
  int **ipartbound;
  struct BRANCH **keen;
  int *keenfc;
  int *keenfcangle;
  int fcqsplit;
  int nsplit;
  int ctfc;
  struct BRANCH **kids;
  struct LEAF *el;
  struct CELL *cl;
} BRANCH;

/* per-leaf geometry (strdata.h:150-168): outward unit normals and tangents per face,
 * volume, sub-face areas Area[f][k], centre, nodes, faces */
typedef struct CELL {
  double *nx;
  double *ny;
  double *nz;
  double *nxt1;
  double *nyt1;
  double *nzt1;
  double *nxt2;
  double *nyt2;
  double *nzt2;
  double Vol;
  double **Area;
  double xc;
  double yc;
  double zc;
  struct ND *nd;
  struct FC *fc;
} CELL;

/* the forest of one rank (strdata.h:170-197): partition buffers, roots, leaf list */
typedef struct FOREST {
  int partitioned;
  int ntrees;
  int ndepth;
  int nleaves;
  int *lleaves;
  int pleaves;
  int tleaves;
  int nleavestmp;
  int ngnd;
  int **proxlfc1;
  int **proxlfc2;
  int **proxtags;
  int **proxdrys;
  int *nprox;
  int **bufflfc1;
  int **bufflfc2;
  int **bufftags;
  int **buffdrys;
  int *nbuff;
  struct TREE **drys;
  struct BRANCH *locl;   /* head of the leaf list */
  struct BRANCH *glob;   /* the node outside the domain (level 0) */
  double *vertex;
  double *vertey;
  double *vertez;
} FOREST;

/* material parameters (strdata.h:199-210) */
typedef struct PAR {
  double matervisc[10];  /* [0]: dynamic viscosity mu */
  int bcoverwrt[10];
  double materpinf[10];  /* [0]: stiffened-gas p_inf */
  double matermush[10];
  double matergama[10];  /* [0]: gamma */
  double materrini[10];
  double materpini[10];
  double u_init[10];
  double v_init[10];
  double w_init[10];
} PAR;

/* run control (strdata.h:212-358).  The hot path reads neq, model, ns, ischeme, iprimtv,
 * limiter, ymin, eos, turb and dt; the rest is the reference's driver state. */
typedef struct CON {
  FV_MPI_COMM comm;
  int level;
  int ns;
  int turb;
  int i_case;
  int iter;
  int nprimitiv;
  int geo_adapth;
  double ymin;
  double r;
  double u;
  double v;
  double w;
  double e;
  double p;
  double g;
  double **amat;
  double **bmat;
  double **cmat;
  double **xmat;
  double *avec;
  double *bvec;
  double *cvec;
  int model;
  int multimat;
  int neqvf;
  int nstep;
  double dt;
  double dt0;
  int exstep;
  int ststep;
  int adapth;
  int adapthstep;
  int start;
  int translator;
  double mscale;
  int splitmode;
  double nxquad, nyquad, nzquad;
  int eos;
  int verbose;
  double RE;
  int halfdt;
  int limiter;
  int istep;
  int tstep;
  int tstep0;
  double time;
  double time0;
  double timetot;
  int rank;
  int size;
  int hmaxlvl;
  double Mach;
  int neq;
  int nueq;
  int nfleq;
  int ndiat;
  int nreac;
  int nlimit;
  char *casename;
  char *runname;
  double maxM1;
  double maxM2;
  double maxM3;
  double maxM4;
  double maxM5;
  int cnamelength;
  char *filend;
  char *fileel;
  char *filecon;
  char *filebr1;
  char *filebr2;
  char *filebr3;
  char *filebr4;
  char *filebr5;
  char *filesor;
  char *filev1;
  char *filev2;
  double stepcomptime;
  double stepcommtime;
  double stepadaptime;
  double timewatch0[20];
  double timewatcht[20];
  double timewatchtall[20];
  int rhsnan;
  int spadr[10];
  int sdadr[10];
  int readr[10];
  double stats[5][10][50][50];
  double statsall[5][10][50][50];
  double stycrd[50];
  double stzcrd[50];
  double stycrdall[50];
  double stzcrdall[50];
  double srcmax;
  double srcmin;
  int nmconseq;
  int neqmass;
  int neqmoment;
  int neqenergy;
  int neq_temp;
  int *eqtypn;
  int *eqtypi;
  double *limitedvars;
  double umax;
  int ischeme;
  int irecons;
  int iprimtv;
  int method;
  int ctype;
  int nindm;
  int idmtp[10];
  int idmmt[10];
  double dmpar[10][10];
} CON;

/* face-state scratch of the reference (strdata.h:360-402) */
typedef struct SOL {
  double r;
  double u;
  double v;
  double w;
  double e;
  double e_internal;
  double p_hydro;
  double u_temp;
  double v_temp;
  double w_temp;
  double **st;
  double c;
  double **flux;
  double *vec;
  double *wec;
  double ux;
  double uy;
  double uz;
  double vx;
  double vy;
  double vz;
  double wx;
  double wy;
  double wz;
  double u_vel;
  double v_vel;
  double w_vel;
  double mu_turb;
  double ymass;
  double vol_frac;
  double r_lm;
  double r_gas;
} SOL;

/* the run (strdata.h:405-540): halo buffers, scratch of the face loops, the control,
 * material and forest structs, timers */
typedef struct RUN {
  double *SendS;
  double *RecvS;
  double *SendSG;
  double *RecvSG;
  double *SendSF;
  double *RecvSF;
  double ***val;
  double ****val4;
  int counter;
  struct SOL *solS;
  struct SOL *solS1;
  struct SOL *sol;
  struct SOL *sol1;
  struct CON *con;
  struct INFO *info;     /* unused by the reference; the GPU binding's per-RUN state */
  struct PAR *par;
  struct FOREST *topo;
  double *AUX_S;
  double *vec_reconstruct;
  double **tr2;
  double *delta0;
  double *delta1;
  double **delta1_face;
  double **delta0_face;
  double **vdelta1;
  double *delta2;
  double *deltal;
  double *deltat;
  double *gradlt;
  double **gradlt_face;
  double *vecl;
  double *wecl;
  double *fluxs;
  double **vecx0;
  double **vecph;
  double **veccf;
  double *xx;
  double **flux_adv;
  double **flux_viscous;
  double *flux_HLLC;
  double **flux_mc;
  double *wec_temp;
  double *vec_temp;
  int flag_ifs2;
  double r_S;
  double e_S;
  double *Us_temp;
  double *tensor_temp;
  double *vel_temp;
  double **wec_face;
  double **wec_face_halfdt;
  int flag_exp;
  int exp_debug;
  int *exp;
  int *exp_part;
  double *exp_RHS;
  double **exp_flux;
  double **exp_fl_hllc;
  int **exp_hllc;
  int **exp_region;
  double **exp_Lar2;
  double **exp_Rar2;
  double **exp_grad;
  double **exp_grad_a2;
  double **exp_grad_r2;
  int deb;
  int adapt_flag_sensor;
  double cputime;
  double stepcputime;
  double Ttot;
  double Tmeshadapt;
  double Tadv;
  double T2nd;
  double Tdist;
  double Tdelta;
  double Trec;
  double T2ndrelax;
  double T2ndcom;
  double T2ndface;
  double Tdist_sum;
  double Tdelta_sum;
  double Trec_sum;
  double T2ndrelax_sum;
  double T2ndcom_sum;
  double T2ndface_sum;
  double Trhs;
  double Trelax;
  double Tcomm;
  double Tinterfloc;
  double Tepx;
  double stepcputime_sum;
  double Ttot_sum;
  double Tmeshadapt_sum;
  double Tadv_sum;
  double T2nd_sum;
  double Trhs_sum;
  double Trelax_sum;
  double Tcomm_sum;
  double Tinterfloc_sum;
  double Tepx_sum;
  int tot_leaves;
} RUN;

/* Reference hot-path entry points (strdata.h:642,654,656,675). */
void RHS(struct RUN *run);
void RK(int irk, RUN *run);
void faceinterp(RUN *run);
void gradcalc(struct RUN *run);

#ifdef __cplusplus
}
#endif

#endif /* FV_STRDATA_H */
