/*
 * oracle.h -- data layout of the CPU oracle for the visionary-sim substep.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this code, and only as the checker or the timed
 * CPU baseline. The product path never calls it.
 *
 * The same source is compiled twice: REAL=float gives the f32 restatement
 * that GPU parity is judged against; REAL=double gives the f64 twin used to
 * state the error of both against (nearly) exact arithmetic. Scalar inputs
 * arrive as double and are rounded to REAL where the reference stores them in
 * an f32 uniform or storage buffer, so the f32 build sees exactly the numbers
 * the reference's kernels see.
 */
#ifndef VSIM_ORACLE_H
#define VSIM_ORACLE_H

#include <stdint.h>

#ifndef REAL
#define REAL float
#endif

/* One lattice body: CellParams (cells.wgsl:18-31) + NodeParams
 * (state.wgsl:24-33) + the lattice origin the splat sampler reads
 * (sample.wgsl:17-21). */
typedef struct {
    uint32_t cell_dims[3];
    uint32_t node_offset;
    uint32_t node_dims[3];
    uint32_t node_count;
    double origin[3];
    double h, alpha_dev, alpha_hyd, gamma, relaxation, max_stretch;
    double friction, damping;
} o_body;

/* One scene: shared node range of every body, the colour-grouped cell list
 * (elasticCells.js:143-163), per-frame uniforms (stepper.js:78-89), contact
 * records and grid (bodyContact.js:69-127), and the kinematic drive
 * (kinematicDrive.js:155-202). State is SoA. */
typedef struct {
    int32_t n_nodes, n_bodies;
    REAL *px, *py, *pz;          /* nodePos  */
    REAL *qx, *qy, *qz;          /* nodePrev */
    REAL *vx, *vy, *vz;          /* nodeVel  */
    const REAL *w;               /* nodeInvMass as the GPU sees it */
    const o_body *bodies;
    const uint32_t *cells;       /* (body, linear cell) pairs, colour-grouped */
    int32_t color_start[8], color_count[8];
    double dt, gravity[3], tank_min[3], tank_max[3];
    /* contact */
    int32_t n_surf, contact;
    const uint32_t *s_node, *s_body;
    const REAL *s_radius;
    double grid_origin[3], bucket_size, gain;
    uint32_t grid_dims[3], bucket_slots;
    /* kinematic drive: law 0 = hard, 1 = compliant */
    int32_t n_driven, drive_law;
    const uint32_t *d_node;
    const REAL *d_off;           /* 3 per record */
    double target[3], target_vel[3], response_time;
    /* optional contact outputs (NULL to skip) */
    uint32_t *key_out;           /* bucket of each surface record */
    int32_t *pair_count;         /* pairs passing the reach test, per record */
    uint32_t *pairs;             /* mate record ids, pair_cap per record */
    int32_t pair_cap;
    int32_t literal;             /* 1: elastic pass in project.wgsl's literal
                                    formulas and contact sums in collide.wgsl's
                                    order; 0: the GPU's fused forms and order */
    int32_t pad1;
    double sphere[4];            /* sphere collider: centre, radius (<= 0: none) */
} o_scene;

/* Splats of one scene; each splat rides one body's lattice. Outputs are SoA:
 * cov holds 6 planes of n values (xx, xy, xz, yy, yz, zz). */
typedef struct {
    int32_t n, pad0;
    const uint32_t *body;
    const REAL *mx, *my, *mz;    /* rest means */
    const REAL *quat;            /* xyzw, 4 per splat */
    const REAL *sx, *sy, *sz;    /* axis standard deviations */
    REAL *ox, *oy, *oz;          /* deformed means */
    REAL *cov;                   /* 6 * n */
    REAL *rot;                   /* engine axes (NULL to skip): 4 planes xyzw */
    REAL *axis;                  /* and 3 planes of axis lengths */
} o_splats;

#endif
