/*
 * sv_oracle.h -- FP64 CPU restatement of the SUNSET LABFM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in the product (sunset-viscoelastic_amd/) links or calls it.
 *
 * Every routine restates the reference algorithm (file:line given per function
 * in sv_oracle.c) with the reference's loop structure: one OpenMP sweep per
 * derivative call, ghost ("mirror") nodes stored explicitly at ids >= npfb,
 * neighbour lists ij_link[i*nplink + k] sorted by distance.  Indices are
 * 0-based here (the reference is 1-based).
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   or_create_ghosts     -- bit-exact vs the flang-built reference (oracle/_ref)
 *   or_find_neighbours   -- bit-exact vs oracle/_ref and vs brute force O(N^2)
 *   or_wab/monomials/abfs-- closed forms (Wendland C2, Hermite recurrence, Taylor)
 *   or_svd_solve         -- LAPACK (numpy) on random systems; oracle/_ref weights
 *   or_labf_weights etc. -- oracle/_ref to 1e-10; moment conditions; polynomial
 *                           exactness
 *   stencil operators    -- dense matvec identity; oracle/_ref derivatives
 *   or_calc_all_rhs (LC) -- oracle/_ref fields after N RK3 steps
 *   or_calc_all_rhs (DI) -- closed forms only (reference DI reads uninitialised
 *                           memory, rhs_di.F90:188-190,421-424); parity for the
 *                           DI form is pinned by relaxation/shear closed forms
 *                           and LC<->DI consistency, not by the reference.
 *   or_boundary_weights, boundary branches (LC) -- oracle/_ref (weights, steps;
 *                           walls, inflow types, ramp, curved walls)
 *   boundary branches (DI) -- the LC branch at c = I; wall polymer diffusion by
 *                           closed forms (normal quartic, tangential cosine)
 *   or_step_rk3          -- oracle/_ref (LC); Butcher order conditions
 *   or_*_tlm (DI, fixed dt) -- base trajectory bit-identical to or_run; tangent
 *                           vs central finite differences of or_run
 */
#ifndef SV_ORACLE_H
#define SV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_FORM_DI = 0, OR_FORM_LC = 1 };
/* time-step modes: PID control (set_tstep_PID, step.F90:624-682), a fixed dt,
 * or dt = min(dt_parabolic, dt_cfl) each step (the non-PID line of set_tstep,
 * step.F90:610-611) */
enum { OR_DT_PID = 0, OR_DT_FIXED = 1, OR_DT_CFL = 2 };

/* Geometry + stencils: the state the reference keeps in common_vars.F90:51-95 */
typedef struct {
  int npfb, np, nplink, cap;     /* real nodes, real+ghost nodes, max links, alloc */
  double *x, *y, *h, *hs, *s;    /* [cap] positions, h, h_small, spacing */
  int *node_type, *irel, *vrel;  /* [cap]; irel/vrel valid for ghosts */
  double xmin, xmax, ymin, ymax, smax, h0, h2;
  int bc[4];                     /* xL xU yL yU: 0 none 1 periodic 2 symmetric 3 no-slip */
  int morder, forder;
  int *ij_count, *ij_count_small, *ij_link; /* [npfb], [npfb], [npfb*nplink] */
  double *w_grad;                /* [npfb][nplink][2] */
  double *w_lap, *w_hyp;         /* [npfb][nplink] */
  double *gsum, *lsum, *hsum, *vol; /* [npfb][2], [npfb], [npfb], [npfb] */
  double L_char;
  /* Boundaries (setup_domain.F90:158-202, :314-326): node types 0 wall,
   * 1 inflow, 2 outflow, each followed in index order by its four FD rows
   * (types -1..-4) along the normal.  Set by or_set_boundaries. */
  int nb;
  int *blist;                    /* [nb] boundary nodes, ascending */
  int *fdp;                      /* [cap] FD parent of a row node, -1 otherwise */
  double *rnx, *rny;             /* [cap] unit normal into the fluid (ghosts: parent's) */
  /* ij_wb_grad2(2,:,jj) and ij_wb_grad2_sum(2,jj): the tangential second
   * derivative of boundary node jj (labf.F90:755, :428-443), divided by
   * L_char^2 in or_calc_labf_sums; the Laplacian of c on boundary nodes in the
   * direct form (calc_laplacian_transverse_only_on_bound, derivatives.F90:306-358) */
  double *wbt, *wbtsum;          /* [nb][nplink], [nb] */
} or_domain;

typedef struct {
  double L_char, U_char, Re, Wi, beta, epsPTT, Ma, Mdiff;
  double grav[3], drive[3];
  double kol_amp, kol_k;         /* f_x = kol_amp*sin(kol_k*y); not in reference LC/DI */
  int form, dt_mode;
  /* derived as load_data.F90:132-148 */
  double rho_char, visc_total, visc_solvent, visc_polymeric, lambda, csq;
  /* inflow boundaries (load_data.F90:152): type 0 soft, 1 hard, 2 inflow-outflow;
   * velocity control 0 fixed (the initial u), 1 ramp u_start -> u_end over ramp_time */
  int inflow_type, u_control;
  double u_start, u_end, ramp_time;
} or_phys;

typedef struct {
  int np;
  double *ro, *rou, *rov, *u, *v;            /* [np] */
  double *pxx, *pxy, *pyy;                   /* log-conformation psi [np] (LC) */
  double *cxx, *cxy, *cyy;                   /* conformation [np] */
  double dt, time, dt_cfl, dt_par;
  double emax_np1, emax_n, emax_nm1;
  int itime;
  double *u_in;                  /* [np] u_inflow_local of inflow nodes (setup_flow.F90:48-51, :130-137) */
} or_state;

/* ---- domain ---- */
int  or_domain_init(or_domain *d, int npfb, const double *x, const double *y,
                    const double *s, const double *h, const double *hs,
                    const int *node_type, double xmin, double xmax, double ymin,
                    double ymax, const int bc[4], double smax, int morder,
                    int forder, double L_char);
void or_domain_free(or_domain *d);
int  or_create_ghosts(or_domain *d);
int  or_find_neighbours(or_domain *d, int nplink);
int  or_calc_labf_weights(or_domain *d);     /* grad, lap, hyp(raw), vol */
void or_calc_labf_sums(or_domain *d);
void or_filter_coefficients(or_domain *d);
/* Boundary nodes (after or_create_ghosts): normals for the npfb real nodes;
 * checks the FD-row layout, builds the boundary list and FD parents.  -1 on a
 * malformed layout. */
int  or_set_boundaries(or_domain *d, const double *rnx, const double *rny);
/* calc_boundary_weights (labf.F90:505-1039): between or_calc_labf_weights and
 * or_calc_labf_sums */
int  or_boundary_weights(or_domain *d);
/* per-node order map (nullable): n-system size chosen per node (config C3) */
int  or_calc_labf_weights_ordermap(or_domain *d, const int *order_of_node,
                                   int nthreads);

/* Checker set-up for full-size node sets (lists and weights computed
 * elsewhere, e.g. downloaded from the device): 0-based lists [npfb][nplink]
 * (after or_create_ghosts), finished weights in the reference layout, and a
 * weight solve of listed nodes only (per-node order nullable). */
int  or_set_stencil(or_domain *d, int nplink, const int *cnt, const int *cnt_small,
                    const int *link);
int  or_set_weights(or_domain *d, const double *w_grad, const double *w_lap,
                    const double *w_hyp, const double *gsum, const double *lsum,
                    const double *hsum, const double *vol);
int  or_weights_subset(or_domain *d, const int *nodes, int count, const int *order_of_node);

/* ---- stencil adaptation (sv_oracle_adapt.c) ---- */
/* Golub-Reinsch SVD in svdlib.F90's operation order: U (n x n column-major,
 * in: A) <- U, W singular values, V, rv1 work; -1 without convergence */
int  or_svd_gr(int n, double *U, double *W, double *V, double *rv1);
void or_svd_gr_backsub(int n, const double *U, const double *W, const double *V, double *b,
                       double *tmp);
int  or_svd_solve_gr(int n, const double *A, double *b);   /* svd_solve */
/* adapt_stencils + adapt_stencils_main (labf.F90:1079-1721) after
 * or_find_neighbours: h, h_small, the lists and counts of real nodes updated;
 * reductions (nullable, 2*npfb) the accepted 2% steps per node and pass */
int  or_adapt_stencils(or_domain *d, int *reductions);
/* the acceptance checks of given stencils (h or h_small): residual measure
 * over its tolerance and the first failing amplification wavenumber */
int  or_adapt_check(or_domain *d, const double *h_sel, int pass, double *ratio, int *amp);

/* ---- kernels / bases ---- */
double or_wab(double q, double h2);
void   or_set_wab_pow4(int mode);   /* 0: libm pow(t, 4.0); 1: correctly rounded */
void   or_monomials(int order, double x, double y, double *out);
void   or_abfs(int order, double x, double y, double *out);
int    or_svd_solve(int n, const double *A, double *b);  /* A column-major n*n */
void   or_eigens2(double a, double b, double d, double l[2], double R[4]);
void   or_c_from_psi(double pxx, double pxy, double pyy, double *cxx, double *cxy,
                     double *cyy);
void   or_psi_from_c(double cxx, double cxy, double cyy, double *pxx, double *pxy,
                     double *pyy);

/* ---- stencil operators (phi sized np, outputs sized npfb) ---- */
void or_calc_gradient(const or_domain *d, const double *phi, double *gx, double *gy);
void or_calc_laplacian(const or_domain *d, const double *phi, double *lap);
void or_calc_laplacian_transverse(const or_domain *d, const double *phi, double *lap);
void or_calc_gradient_only(const or_domain *d, int dim, const double *phi, double *g);
void or_calc_filtered_var(const or_domain *d, double *phi);
void or_calc_filter_term(const or_domain *d, const double *phi, double *out);

/* ---- physics ---- */
void or_phys_derive(or_phys *p);
int  or_state_alloc(or_state *st, int np);
void or_state_free(or_state *st);
/* Initial conditions as hardcode_initial_conditions/initial_solution
 * (setup_flow.F90:24-173): ro,u,v,c given on real nodes; psi=log(c). */
void or_set_initial(const or_domain *d, const or_phys *p, or_state *st,
                    const double *ro, const double *u, const double *v,
                    const double *cxx, const double *cxy, const double *cyy);
void or_reapply_mirror_bcs(const or_domain *d, const or_phys *p, or_state *st);
/* calc_all_rhs: outputs rhs_* [npfb]; iRKstep as in rhs_lc.F90:503 */
void or_calc_all_rhs(const or_domain *d, const or_phys *p, const or_state *st,
                     int iRKstep, double *r_ro, double *r_rou, double *r_rov,
                     double *r_xx, double *r_xy, double *r_yy);
void or_filter_variables(const or_domain *d, const or_phys *p, or_state *st);
void or_set_tstep(const or_domain *d, const or_phys *p, or_state *st);
void or_set_tstep_pid(or_state *st);
void or_step_rk3(const or_domain *d, const or_phys *p, or_state *st);
/* time loop as sunset.F90:75-106 (set_tstep, PID unless fixed dt, step, itime++) */
void or_run(const or_domain *d, const or_phys *p, or_state *st, int nsteps);

int  or_num_threads(void);
void or_set_num_threads(int n);

/* ---- diagnostics (statistics.F90:305-654) ---- */
void   or_diag_mass_energy(const or_domain *d, const or_phys *p, const or_state *st,
                           double out[3]);
void   or_diag_conf(const or_domain *d, const or_state *st, double out[6]);
void   or_diag_kolmogorov(const or_domain *d, const or_phys *p, const or_state *st,
                          double out[5]);
double or_diag_tg(const or_domain *d, const or_phys *p, const or_state *st);

/* ---- tangent-linear model of the DI step at a fixed dt (sv_oracle_tlm.c) ---- */
typedef struct {
  int np;
  double *ro, *rou, *rov, *u, *v, *cxx, *cxy, *cyy;   /* [np] tangents */
  double dWi, dbeta, deps;   /* tangent of the parameters Wi, beta, epsPTT (0 = state only) */
} or_tangent;
int  or_tangent_alloc(or_tangent *t, int np);
void or_tangent_free(or_tangent *t);
void or_tangent_set(const or_domain *d, const or_state *st, or_tangent *t, const double *dro,
                    const double *drou, const double *drov, const double *dcxx,
                    const double *dcxy, const double *dcyy);
void or_calc_all_rhs_di_tlm(const or_domain *d, const or_phys *p, const or_state *st,
                            const or_tangent *t, double *const dr[6]);
int  or_step_rk3_di_tlm(const or_domain *d, const or_phys *p, or_state *st, or_tangent *t);
double or_loss_trace(const or_domain *d, const or_state *st);
int  or_rollout_tlm(const or_domain *d, const or_phys *p, or_state *st, or_tangent *t,
                    int nsteps, double *loss, double *dloss);

/* ---- planted mistakes (CPU tests only): or_set_plant(code) switches one slip
 *      into or_calc_all_rhs until it is set back to OR_PLANT_NONE ---------- */
enum {
  OR_PLANT_NONE = 0,
  OR_PLANT_DI_ADV_SWAP = 1,       /* DI advection u d/dy c + v d/dx c          */
  OR_PLANT_DI_XONLY_GRAD = 2,     /* DI: d/dy cxx = d/dx cyy = 0 (the reference's
                                     calc_gradient_only use, rhs_di.F90:188-190) */
  OR_PLANT_DI_ADV_SIGN = 3,       /* DI advection +u.grad c                    */
  OR_PLANT_MDIFF_DROP = 4,        /* DI polymer diffusion left out             */
  OR_PLANT_MDIFF_XX = 5,          /* DI diffusion of cxx used for every component */
  OR_PLANT_KOL_RHO = 6,           /* Kolmogorov forcing times rho              */
  OR_PLANT_KOL_SIGN = 7,          /* Kolmogorov forcing with the opposite sign */
  OR_PLANT_KOL_X = 8,             /* Kolmogorov forcing sin(k x)               */
  OR_PLANT_KOL_V = 9,             /* Kolmogorov forcing on the v momentum      */
  OR_PLANT_DIVC_TRANSPOSED = 10   /* polymer stress divergence d/dy cxx + d/dx cxy */
};
void or_set_plant(int code);

#ifdef __cplusplus
}
#endif
#endif
