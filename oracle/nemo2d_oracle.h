/* CPU oracle for the NEMOLite2D time step -- test infrastructure only.
 * A C restatement of original/nemolite2d.f90 (the reference's serial model);
 * see nemo2d_oracle.c for how it is pinned. */
#ifndef NEMO2D_ORACLE_H
#define NEMO2D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* &namctl physics/time parameters, original/nemolite2d.f90:100-118. */
typedef struct {
    double dx, dy, dep_const, rdt, cbfr, visc;
} o2d_params;

typedef struct o2d o2d;

o2d *o2d_create(int jpi, int jpj);
void o2d_destroy(o2d *m);
void o2d_set_params(o2d *m, const o2d_params *p);
int o2d_jpi(const o2d *m);
int o2d_jpj(const o2d *m);
double *o2d_array(o2d *m, const char *name, int bounds[4]);
int *o2d_pt(o2d *m);
void o2d_grid_default(o2d *m);
void o2d_grid_areas(o2d *m);
void o2d_initialise(o2d *m);
void o2d_continuity(o2d *m);
void o2d_momentum_u(o2d *m);
void o2d_momentum_v(o2d *m);
double o2d_tide(double rtime);
void o2d_set_tide(o2d *m, double amp, double period_hours);
void o2d_bc_ssh(o2d *m, double rtime);
void o2d_bc_solid_u(o2d *m);
void o2d_bc_solid_v(o2d *m);
void o2d_bc_flather_u(o2d *m);
void o2d_bc_flather_v(o2d *m);
void o2d_bc(o2d *m, double rtime);
void o2d_next(o2d *m);
void o2d_step(o2d *m, int istp);
void o2d_checksums(o2d *m, double out[3]);
double o2d_run(o2d *m, int nit000, int nitend, double *sums);

#ifdef __cplusplus
}
#endif
#endif
