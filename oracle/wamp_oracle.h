/*
 * wamp_oracle.h -- CPU ORACLE (test infrastructure only; see wamp_oracle.c).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 * The same source also builds as liborc_count.so (opcount.h), which counts the
 * floating-point operations of a step by kind.
 */
#ifndef WAMP_ORACLE_H
#define WAMP_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_eos {
    double g1, p1;   /* phase 1 (gamma, p_inf), SC:40-41 */
    double g2, p2;   /* phase 2, SC:42-43 */
} orc_eos;

typedef struct orc_params {
    int nx, ny, g;        /* interior cells and ghost layers, SC:26-27 */
    double dx, dy;        /* SC:31-32 */
    orc_eos eos;
    int base;             /* 0 MUSCL, 1 MP5 */
    int shear;            /* 0 none, 1 central4, 2 central6 (Ducros-gated) */
    int vf_thinc;         /* 0 gated, 1 always */
    int face_pp;          /* 0 simple (SC), 1 Wong + pressure floor (M10) */
    int repair;           /* USE_EMERGENCY_REPAIR, SC:37 */
    int bc[4];            /* x-lo, x-hi, y-lo, y-hi: 0 zero-gradient, 1 slip wall */
    double cfl;           /* SC:34 */
} orc_params;

typedef struct orc_s6 { double v0, v1, v2, v3, v4, v5; } orc_s6;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_params *p);
void orc_destroy(orc_ctx *c);
double *orc_array(orc_ctx *c, int which);
void orc_set_threads(int n);
int orc_max_threads(void);

void orc_prepare(orc_ctx *c);
void orc_apply_bc(orc_ctx *c);
void orc_update_prims(orc_ctx *c);
void orc_enforce_admissible(orc_ctx *c);
void orc_compute_fluxes(orc_ctx *c);
void orc_limit_fluxes(orc_ctx *c, double dt);
void orc_rk3_update(orc_ctx *c, double alpha, double beta, double dt);
void orc_copy_old(orc_ctx *c);
double orc_get_dt(orc_ctx *c);
void orc_step(orc_ctx *c, double dt);

/* Test hook (no counterpart in the reference): while set, orc_step calls
   fn(user, stage) for stage 0, 1, 2 in place of compute_fluxes +
   compute_low_fluxes, so a test can drive the stage sequence (limiter,
   SSP-RK3 combine, repair, prims) with face fluxes of its choosing.  NULL
   restores the flux kernels. */
typedef void (*orc_flux_hook)(void *user, int stage);
void orc_set_flux_hook(orc_ctx *c, orc_flux_hook fn, void *user);
void orc_run(orc_ctx *c, int nsteps, double *dts);

/* scalar helpers (exported for known-answer tests) */
double orc_pow(double x, double y);
double orc_tanh(double x);
void orc_mix_gamma_pibig(const orc_eos *e, double a1, double *gb, double *pb);
void orc_mix_gamma_pinf(const orc_eos *e, double a1, double *g, double *pi);
double orc_sound_speed(const orc_eos *e, double rho, double p, double a1);
double orc_prim_to_energy(const orc_eos *e, double ar1, double ar2, double u,
                          double v, double p, double a1);
double orc_clamp01_pp(double a1);
double orc_pp_theta_lower(double base, double high, double floor_);
orc_s6 orc_pp_limit_reconstructed_state(const orc_eos *e, orc_s6 high, orc_s6 base);
orc_s6 orc_pp_floor_pressure_state(const orc_eos *e, orc_s6 s);
double orc_rho_ctilde2_cons(const orc_eos *e, double m1, double m2, double mx,
                            double my, double E, double a1_in);
double orc_theta_cons_stage(const orc_eos *e, const double l[6], const double h[6], int stage);
double orc_cell_entropy(const orc_eos *e, double ar1, double ar2, double p, double a1);
double orc_cator2_val(double em2, double em1, double e0, double ep1, double ep2);
double orc_thinc_ul(double vm1, double v0, double v1);
double orc_thinc_ur(double v0, double v1, double v2);
double orc_mp5_ul(double vm2, double vm1, double v0, double vp1, double vp2);
double orc_mp5_ur(double vm1, double v0, double vp1, double vp2, double vp3);
double orc_central6_avg(double vm2, double vm1, double v0, double vp1, double vp2, double vp3);
double orc_central4(double vm1, double v0, double vp1, double vp2);
double orc_muscl_char_ul(double vm1, double v0, double vp1, double vp2);
double orc_muscl_char_ur(double vm1, double v0, double vp1, double vp2);
orc_s6 orc_to_char_x(const double u[6], double c, double rho, double ar1,
                     double ar2, double u_avg, double v_avg);
orc_s6 orc_to_phys_x(orc_s6 w, double c, double rho, double ar1, double ar2,
                     double u_avg, double v_avg);
orc_s6 orc_to_char_y(const double u[6], double c, double rho, double ar1,
                     double ar2, double u_avg, double v_avg);
orc_s6 orc_to_phys_y(orc_s6 w, double c, double rho, double ar1, double ar2,
                     double u_avg, double v_avg);
orc_s6 orc_hllc_x(const orc_eos *e, const double L[6], const double R[6]);
orc_s6 orc_hllc_y(const orc_eos *e, const double L[6], const double R[6]);
double orc_hllc_x_src(const orc_eos *e, const double L[5], const double R[5]);
double orc_hllc_y_src(const orc_eos *e, const double L[5], const double R[5]);

#ifdef __cplusplus
}
#endif
#endif
