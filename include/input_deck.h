/*
 * input_deck.h -- the key = value input deck of the solver.
 *
 * Drop-in for the reference's input_deck.h (input_deck.h:1-132): the two
 * enums keep their enumerator order and values, InputDeck keeps every field
 * in the reference's order, and the five functions keep their signatures.
 * Fields added by this build sit after the reference's last field, so code
 * compiled against the reference layout that only holds InputDeck pointers
 * returned by input_deck_create keeps working.
 */
#ifndef INPUT_DECK_H
#define INPUT_DECK_H

#include <stdbool.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* input_deck.h:8-27 */
typedef enum {
    KELVIN_HELMHOLTZ,
    BLAST_WAVE,
    RAYLEIGH_TAYLOR,
    RIEMANN_2D,
    STALLED_SHOCK,
    STELLAR_CONVECTION,
    STELLAR_PROFILE,
    GALAXY_DISK,
    UNIFORM_DISK_TEST,
    WD_MERGER,
    SUPERNOVA,
    KELVIN_HELMHOLTZ_3D,
    BLAST_WAVE_3D,
    TURBULENCE_3D,
    SEDOV_BLAST_3D,
    STELLAR_PROFILE_3D,
    BINARY_STAR_3D,
    STELLAR_OCTANT_EXPLOSION_3D
} ProblemType;

/* input_deck.h:30-34 */
typedef enum { BC_PERIODIC, BC_REFLECTIVE, BC_OUTFLOW } BCType;

/* Poisson solve modes added by this build. */
typedef enum {
    POISSON_FP64_SOR = 0,   /* plain red-black SOR in FP64 (the reference) */
    POISSON_FP32_DC = 1,    /* FP64 defect correction, FP32 red-black sweeps */
    POISSON_NN_FP32_DC = 2  /* learned initial guess (kept only when its
                               residual beats the warm start), then FP32_DC */
} PoissonMode;

/* Inner solver of the FP32 defect correction, added by this build. */
typedef enum {
    DC_INNER_SOR = 0,       /* red-black SOR sweeps on the solver's grid */
    DC_INNER_MULTIGRID = 1  /* V-cycles of red-black sweeps over a grid
                               hierarchy (even nx, ny, nz) */
} DcInner;

/* Order of the multipole moment sums, added by this build. */
typedef enum {
    MOMENT_ORDER_TREE = 0,       /* pairwise double-double tree (compensated) */
    MOMENT_ORDER_SEQUENTIAL = 1, /* cell after cell as the reference
                                    (gravity_solver_3d.c:182-198, 226-258):
                                    the same bits as its loops            */
    MOMENT_ORDER_AUTO = 2        /* the default: SEQUENTIAL for plain FP64
                                    SOR (the reference's results), TREE for
                                    the FP32 defect-correction modes        */
} MomentOrder;

/* SOR stopping rules added by this build. */
typedef enum {
    SOR_CRIT_MAX_UPDATE = 0, /* max |phi_GS - phi_old| < tol (the reference) */
    SOR_CRIT_REL_RESID = 1   /* ||4piG rho - L phi||_2 / ||4piG rho||_2 <= tol */
} SorCriterion;

/* input_deck.h:37-124, then the fields this build adds. */
typedef struct {
    int nx, ny, nz;
    double xmin, xmax;
    double ymin, ymax;
    double zmin, zmax;

    double t_end;
    double output_dt;
    double cfl;

    double gamma;
    double G_const;

    ProblemType problem;

    BCType bc_xmin, bc_xmax;
    BCType bc_ymin, bc_ymax;
    BCType bc_zmin, bc_zmax;

    double G_M_PNS;
    double L_nu;
    double g_strength;
    double L_heat;

    char output_dir[256];

    bool restart;
    char restart_file[256];
    int restart_step;
    double restart_time;

    char star1_profile[256];
    char star2_profile[256];
    double star1_x, star1_y, star1_z;
    double star2_x, star2_y, star2_z;
    double star1_vx, star1_vy, star1_vz;
    double star2_vx, star2_vy, star2_vz;
    double binary_separation;
    double binary_orbital_velocity;
    bool binary_auto_orbit;

    char stellar_profile[256];
    double explosion_energy;
    double explosion_radius;
    char explosion_type[32];

    bool enable_perturbations;
    int perturbation_seed;

    double perturb_velocity_amplitude;
    double perturb_velocity_kmin;
    double perturb_velocity_kmax;
    int perturb_velocity_modes;

    double perturb_density_amplitude;
    double perturb_density_kmin;
    double perturb_density_kmax;
    int perturb_density_modes;

    bool enable_sph_harmonic_mode;
    int sph_harmonic_l;
    int sph_harmonic_m;
    double sph_harmonic_amp_rho;
    double sph_harmonic_amp_vel;

    bool enable_torus;
    double torus_major_radius;
    double torus_minor_radius;
    double torus_density;
    double torus_pressure;
    double torus_rotation_velocity;

    /* ---- added by this build (defaults reproduce the reference) ---- */
    int max_steps;            /* 0: run to t_end only                     */
    double sor_omega;         /* 1.5  (gravity_solver_3d.c:78)            */
    int sor_max_iter;         /* 10000 (gravity_solver_3d.c:79)           */
    double sor_tol;           /* 1e-6 (gravity_solver_3d.c:80)            */
    int sor_criterion;        /* SorCriterion, default max update         */
    int multipole_lmax;       /* 2: monopole + quadrupole (the reference) */
    int poisson_mode;         /* PoissonMode, default FP64 SOR            */
    char nn_guess[256];       /* weights file for the learned guess, ""   */
    bool output_enable;       /* false skips CSV/VTK writing              */
    char ic_file[256];        /* binary initial state written by this build */
    int dc_inner;             /* DcInner of the FP32 modes, default SOR   */
    int moment_order;         /* MomentOrder, default AUTO                */
    bool checkpoint;          /* also write HYDCK001 checkpoints at output
                                 times (U, potential, g), default false   */
    char nn_correction[256];  /* weights file for the start correction of
                                 POISSON_NN_FP32_DC, ""                   */
} InputDeck;

/* input_deck.h:127-131 */
InputDeck *input_deck_create(const char *filename);
void input_deck_destroy(InputDeck *deck);
void input_deck_print(const InputDeck *deck);
ProblemType parse_problem_type(const char *str);
BCType parse_bc_type(const char *str);

#ifdef __cplusplus
}
#endif
#endif /* INPUT_DECK_H */
