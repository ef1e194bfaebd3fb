/* enl_oracle.h -- CPU restatement of the reference kernels (test
 * infrastructure only; see enl_oracle.c). */
#ifndef ENL_ORACLE_H
#define ENL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* rhs variants: one per fortran/rhs_*.m4 file (makefile:6 picks one). */
enum {
  ENL_RHS_HYBRID_HIGH_LLF = 0,
  ENL_RHS_HYBRID_HIGH = 1,
  ENL_RHS_HYBRID_LOW_LLF = 2,
  ENL_RHS_HYBRID_LOW = 3,
  ENL_RHS_WENO_HIGH_LLF = 4,
  ENL_RHS_WENO_HIGH = 5,
  ENL_RHS_WENO_LOW_LLF = 6,
  ENL_RHS_WENO_LOW = 7,
  ENL_RHS_DRP = 8
};

/* The Enlightning{constants} block (input.txt:40-59) plus the derived tau_n,
 * tau_o (enlightning.C:218-234) and tag_tolerance (input.txt:4). */
typedef struct {
  double weno_alpha, c, gamma1, p0, T0, c_lr, rho0, R, c_p, T_star_n, T_star_o;
  double mu, muB, kappa, c_v, R_tilde, M, tau_n, tau_o;
  double tag_tolerance[3];
} enl_gas_t;

double orc_wenom5(double, double, double, double, double);
double orc_wenop5(double, double, double, double, double);
double orc_wenom3(double, double, double);
double orc_wenop3(double, double, double);
double orc_sensor_eps(double tol);
int orc_rhs_ghost_need(int variant);
void orc_literals(double *out);

int orc_rhs_variant(int variant, const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                    const int *gcw0, const int *gcw1, const double *xylo, const double *dxy,
                    const double *w_n, const double *H0, double *KK, const double *pres, const double *temp,
                    const double *alpha, const double *kappa, const double *mu, const double *muB,
                    const double *R, const double *T_star_n, const double *T_star_o, const double *tau_n,
                    const double *tau_o, const double *tolerance, const int *level, const double *c,
                    const double *c_lr, const int *nequ, int *switch_map);

void orc_runge_kutta(const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                     const int *gcw0, const int *gcw1, const double *dt, const double *alpha_1,
                     const double *alpha_2, const double *beta, double *w_n, const double *ww,
                     const double *KK, const int *nequ);

void orc_update_state(const double *dxy, const double *xylo, const int *ifirst0, const int *ilast0,
                      const int *ifirst1, const int *ilast1, const int *gcw0, const int *gcw1,
                      const double *w_n, double *pres, double *temp, const double *rho0, const double *c,
                      const double *gamma1, const double *c_p, const double *p0, const double *T0,
                      const double *c_lr, const double *c_v, const double *R_tilde, const double *M,
                      const double *R, const int *nequ);

void orc_tag_cells(const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                   const int *gcw0, const int *gcw1, int *tags, const double *w_n, const double *H0,
                   const int *refine_tag_val, const double *tolerance, const int *nequ);

void orc_fill_physical_bc(int if0, int il0, int if1, int il1, int g0, int g1, const int *touch,
                          const int *reflect, double *w_n, double *pres, double *temp, double *source);

int orc_advance_patch(int variant, int if0, int il0, int if1, int il1, int g, const double *xylo,
                      const double *dxy, const enl_gas_t *gas, const int *reflect, int order,
                      const double *alpha_1, const double *alpha_2, const double *beta, double dt,
                      int nsteps, int level, double *w_n, double *pres, double *temp, double *source,
                      double *ww, double *KK, int tile, int nthreads, double absorb_width);

/* Source parameters (input.txt:20-38); mode 10 PULSE / 20 DRIVEN, type
 * 10 POINT / 20 SINE / 30 LIGHTNING (enlightning.C:55-61). */
typedef struct {
  int mode, type;
  double amp, spread, pos[2], sine_amp, sine_freq, omega;
  int nseg;
  const double *coords;
} orc_source_t;

int orc_add_source(int if0, int il0, int if1, int il1, const double *pxlo, const double *dx,
                   const orc_source_t *src, int iteration, int pulse_time0, double dt, double *source, long sy);
int orc_read_lightning(const char *path, double *coords, int max_segments);
void orc_sample_mics(int if0, int il0, int if1, int il1, const double *pxlo, const double *pxhi, const double *dx,
                     int nmics, const double *mx, const double *my, const double *pres, long sy, double *out,
                     int *hit);

void orc_absorb(int if0, int il0, int if1, int il1, int g0, int g1, const double *pxlo, const double *dx,
                const double *gxlo, const double *gxhi, const int *absorb, double width, const enl_gas_t *gas,
                double *w_n, double *pres, double *temp);

#ifdef __cplusplus
}
#endif
#endif
