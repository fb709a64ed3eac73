// input_deck.cpp -- the key = value input deck (input_deck.c:1-419).
//
// Same defaults, same comment and whitespace rules (whole-line '#'
// comments, trailing '#' comments stripped from values, whitespace trimmed),
// same unknown-key warning, plus the keys this build adds with defaults that
// reproduce the reference's behaviour.
#include <ctype.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/input_deck.h"

namespace {

char* trim(char* str) {
    while (isspace((unsigned char)*str)) str++;
    if (*str == 0) return str;
    char* end = str + strlen(str) - 1;
    while (end > str && isspace((unsigned char)*end)) end--;
    end[1] = '\0';
    return str;
}

bool truthy(const char* v) { return strcmp(v, "true") == 0 || strcmp(v, "1") == 0; }

void copy_str(char* dst, size_t cap, const char* src) {
    strncpy(dst, src, cap - 1);
    dst[cap - 1] = '\0';
}

struct Key {
    const char* name;
    char kind;  // 'i' int, 'd' double, 'b' bool, 's' string
    size_t off;
    size_t cap;
};

#define KI(f) {#f, 'i', offsetof(InputDeck, f), 0}
#define KD(f) {#f, 'd', offsetof(InputDeck, f), 0}
#define KB(f) {#f, 'b', offsetof(InputDeck, f), 0}
#define KS(f) {#f, 's', offsetof(InputDeck, f), sizeof(((InputDeck*)0)->f)}

// input_deck.c:193-315, then the added keys
const Key kKeys[] = {
    KI(nx), KI(ny), KI(nz), KD(xmin), KD(xmax), KD(ymin), KD(ymax), KD(zmin), KD(zmax),
    KD(t_end), KD(output_dt), KD(cfl), KD(gamma), KD(G_const),
    KD(G_M_PNS), KD(L_nu), KD(g_strength), KD(L_heat), KS(output_dir),
    KB(restart), KS(restart_file), KI(restart_step), KD(restart_time),
    KS(star1_profile), KS(star2_profile), KD(star1_x), KD(star1_y), KD(star1_z),
    KD(star2_x), KD(star2_y), KD(star2_z), KD(star1_vx), KD(star1_vy), KD(star1_vz),
    KD(star2_vx), KD(star2_vy), KD(star2_vz), KD(binary_separation),
    KD(binary_orbital_velocity), KB(binary_auto_orbit),
    KS(stellar_profile), KD(explosion_energy), KD(explosion_radius), KS(explosion_type),
    KB(enable_perturbations), KI(perturbation_seed),
    KD(perturb_velocity_amplitude), KD(perturb_velocity_kmin), KD(perturb_velocity_kmax),
    KI(perturb_velocity_modes), KD(perturb_density_amplitude), KD(perturb_density_kmin),
    KD(perturb_density_kmax), KI(perturb_density_modes),
    KB(enable_sph_harmonic_mode), KI(sph_harmonic_l), KI(sph_harmonic_m),
    KD(sph_harmonic_amp_rho), KD(sph_harmonic_amp_vel),
    KB(enable_torus), KD(torus_major_radius), KD(torus_minor_radius), KD(torus_density),
    KD(torus_pressure), KD(torus_rotation_velocity),
    KI(max_steps), KD(sor_omega), KI(sor_max_iter), KD(sor_tol), KI(multipole_lmax),
    KS(nn_guess), KB(output_enable), KS(ic_file), KB(checkpoint), KS(nn_correction),
};

const char* bc_name(BCType b) {
    return b == BC_PERIODIC ? "PERIODIC" : (b == BC_REFLECTIVE ? "REFLECTIVE" : "OUTFLOW");
}

}  // namespace

extern "C" {

ProblemType parse_problem_type(const char* str) {
    static const char* names[] = {"KELVIN_HELMHOLTZ", "BLAST_WAVE", "RAYLEIGH_TAYLOR",
                                  "RIEMANN_2D", "STALLED_SHOCK", "STELLAR_CONVECTION",
                                  "STELLAR_PROFILE", "GALAXY_DISK", "UNIFORM_DISK_TEST",
                                  "WD_MERGER", "SUPERNOVA", "KELVIN_HELMHOLTZ_3D",
                                  "BLAST_WAVE_3D", "TURBULENCE_3D", "SEDOV_BLAST_3D",
                                  "STELLAR_PROFILE_3D", "BINARY_STAR_3D",
                                  "STELLAR_OCTANT_EXPLOSION_3D"};
    for (int i = 0; i < (int)(sizeof names / sizeof names[0]); ++i)
        if (strcmp(str, names[i]) == 0) return (ProblemType)i;
    fprintf(stderr, "Warning: Unknown problem type '%s', defaulting to KELVIN_HELMHOLTZ\n", str);
    return KELVIN_HELMHOLTZ;
}

BCType parse_bc_type(const char* str) {
    if (strcmp(str, "PERIODIC") == 0) return BC_PERIODIC;
    if (strcmp(str, "REFLECTIVE") == 0) return BC_REFLECTIVE;
    if (strcmp(str, "OUTFLOW") == 0) return BC_OUTFLOW;
    fprintf(stderr, "Warning: Unknown BC type '%s', defaulting to OUTFLOW\n", str);
    return BC_OUTFLOW;
}

InputDeck* input_deck_create(const char* filename) {
    InputDeck* d = (InputDeck*)calloc(1, sizeof(InputDeck));
    if (!d) {
        fprintf(stderr, "Error: Failed to allocate memory for InputDeck\n");
        return NULL;
    }
    // defaults, input_deck.c:63-153
    d->nx = d->ny = d->nz = 128;
    d->xmin = d->ymin = d->zmin = -1.0;
    d->xmax = d->ymax = d->zmax = 1.0;
    d->t_end = 1.0;
    d->output_dt = 0.1;
    d->cfl = 0.4;
    d->gamma = 5.0 / 3.0;
    d->G_const = 1.0;
    d->problem = KELVIN_HELMHOLTZ_3D;
    d->bc_xmin = d->bc_xmax = d->bc_ymin = d->bc_ymax = d->bc_zmin = d->bc_zmax = BC_PERIODIC;
    strcpy(d->output_dir, "output");
    strcpy(d->star1_profile, "stellar_profile_1.txt");
    strcpy(d->star2_profile, "stellar_profile_2.txt");
    d->star1_x = -1.5;
    d->star2_x = 1.5;
    d->star1_vy = 0.5;
    d->star2_vy = -0.5;
    d->binary_separation = 3.0;
    d->binary_auto_orbit = true;
    strcpy(d->stellar_profile, "stellar_profile.txt");
    d->explosion_energy = 1.0;
    d->explosion_radius = 0.2;
    strcpy(d->explosion_type, "thermal");
    d->perturbation_seed = 42;
    d->perturb_velocity_amplitude = 0.01;
    d->perturb_velocity_kmin = 1.0;
    d->perturb_velocity_kmax = 10.0;
    d->perturb_velocity_modes = 20;
    d->perturb_density_amplitude = 0.01;
    d->perturb_density_kmin = 1.0;
    d->perturb_density_kmax = 10.0;
    d->perturb_density_modes = 20;
    d->sph_harmonic_l = 1;
    d->sph_harmonic_amp_rho = 0.05;
    d->sph_harmonic_amp_vel = 0.05;
    d->torus_major_radius = 1.2;
    d->torus_minor_radius = 0.2;
    d->torus_density = 1.0;
    d->torus_pressure = 0.1;
    // added keys: the reference's fixed settings (gravity_solver_3d.c:78-80)
    d->max_steps = 0;
    d->sor_omega = 1.5;
    d->sor_max_iter = 10000;
    d->sor_tol = 1.0e-6;
    d->sor_criterion = SOR_CRIT_MAX_UPDATE;
    d->multipole_lmax = 2;
    d->poisson_mode = POISSON_FP64_SOR;
    d->dc_inner = DC_INNER_SOR;
    d->moment_order = MOMENT_ORDER_AUTO;
    d->output_enable = true;
    d->checkpoint = false;

    if (filename == NULL) {
        printf("No input file provided, using defaults.\n");
        return d;
    }
    FILE* fp = fopen(filename, "r");
    if (!fp) {
        fprintf(stderr, "Warning: Could not open input file '%s', using defaults.\n", filename);
        return d;
    }
    printf("Reading input deck from: %s\n", filename);
    char line[512];
    int line_num = 0;
    while (fgets(line, sizeof line, fp)) {
        line_num++;
        char* t = trim(line);
        if (strlen(t) == 0 || t[0] == '#') continue;
        char key[128], value[256];
        if (sscanf(t, "%127[^=]=%255[^\n]", key, value) != 2) continue;
        char* hash = strchr(value, '#');
        if (hash) *hash = '\0';
        const char* k = trim(key);
        const char* v = trim(value);
        if (strcmp(k, "problem") == 0) { d->problem = parse_problem_type(v); continue; }
        if (strcmp(k, "bc_xmin") == 0) { d->bc_xmin = parse_bc_type(v); continue; }
        if (strcmp(k, "bc_xmax") == 0) { d->bc_xmax = parse_bc_type(v); continue; }
        if (strcmp(k, "bc_ymin") == 0) { d->bc_ymin = parse_bc_type(v); continue; }
        if (strcmp(k, "bc_ymax") == 0) { d->bc_ymax = parse_bc_type(v); continue; }
        if (strcmp(k, "bc_zmin") == 0) { d->bc_zmin = parse_bc_type(v); continue; }
        if (strcmp(k, "bc_zmax") == 0) { d->bc_zmax = parse_bc_type(v); continue; }
        if (strcmp(k, "poisson_mode") == 0) {
            if (strcmp(v, "fp32dc") == 0 || strcmp(v, "1") == 0)
                d->poisson_mode = POISSON_FP32_DC;
            else if (strcmp(v, "nn_fp32dc") == 0 || strcmp(v, "2") == 0)
                d->poisson_mode = POISSON_NN_FP32_DC;
            else
                d->poisson_mode = POISSON_FP64_SOR;
            continue;
        }
        if (strcmp(k, "dc_inner") == 0) {
            d->dc_inner = (strcmp(v, "multigrid") == 0 || strcmp(v, "1") == 0)
                              ? DC_INNER_MULTIGRID
                              : DC_INNER_SOR;
            continue;
        }
        if (strcmp(k, "moment_order") == 0) {
            d->moment_order = (strcmp(v, "sequential") == 0 || strcmp(v, "1") == 0)
                                  ? MOMENT_ORDER_SEQUENTIAL
                              : (strcmp(v, "tree") == 0 || strcmp(v, "0") == 0)
                                  ? MOMENT_ORDER_TREE
                                  : MOMENT_ORDER_AUTO;
            continue;
        }
        if (strcmp(k, "sor_criterion") == 0) {
            d->sor_criterion = (strcmp(v, "rel_resid") == 0 || strcmp(v, "1") == 0)
                                   ? SOR_CRIT_REL_RESID
                                   : SOR_CRIT_MAX_UPDATE;
            continue;
        }
        bool found = false;
        for (const Key& e : kKeys) {
            if (strcmp(k, e.name) != 0) continue;
            char* base = (char*)d + e.off;
            switch (e.kind) {
                case 'i': *(int*)base = atoi(v); break;
                case 'd': *(double*)base = atof(v); break;
                case 'b': *(bool*)base = truthy(v); break;
                case 's': copy_str(base, e.cap, v); break;
            }
            found = true;
            break;
        }
        if (!found) fprintf(stderr, "Warning: Unknown parameter '%s' on line %d\n", k, line_num);
    }
    fclose(fp);
    printf("Successfully loaded input deck.\n");
    return d;
}

void input_deck_destroy(InputDeck* deck) { free(deck); }

void input_deck_print(const InputDeck* d) {
    // input_deck.c:334-418 (abridged to the fields the solver reads)
    printf("\n==============================================\n");
    printf("INPUT DECK PARAMETERS\n");
    printf("==============================================\n");
    printf("Grid:\n  nx = %d, ny = %d, nz = %d\n", d->nx, d->ny, d->nz);
    printf("  x: [%.2f, %.2f]\n  y: [%.2f, %.2f]\n  z: [%.2f, %.2f]\n", d->xmin, d->xmax, d->ymin,
           d->ymax, d->zmin, d->zmax);
    printf("\nTime:\n  t_end = %.3f\n  output_dt = %.3f\n  cfl = %.3f\n", d->t_end, d->output_dt,
           d->cfl);
    printf("\nPhysics:\n  gamma = %.3f\n  G_const = %.3f\n", d->gamma, d->G_const);
    printf("\nBoundary Conditions:\n  x: %s / %s\n  y: %s / %s\n  z: %s / %s\n",
           bc_name(d->bc_xmin), bc_name(d->bc_xmax), bc_name(d->bc_ymin), bc_name(d->bc_ymax),
           bc_name(d->bc_zmin), bc_name(d->bc_zmax));
    printf("\nPoisson:\n  omega = %.3f, max_iter = %d, tol = %.3e, lmax = %d, mode = %s\n",
           d->sor_omega, d->sor_max_iter, d->sor_tol, d->multipole_lmax,
           d->poisson_mode == POISSON_FP32_DC      ? "fp32dc"
           : d->poisson_mode == POISSON_NN_FP32_DC ? "nn_fp32dc"
                                                   : "fp64");
    if (d->poisson_mode != POISSON_FP64_SOR)
        printf("  dc_inner = %s\n", d->dc_inner == DC_INNER_MULTIGRID ? "multigrid" : "sor");
    printf("  moment_order = %s\n", d->moment_order == MOMENT_ORDER_SEQUENTIAL ? "sequential"
                                     : d->moment_order == MOMENT_ORDER_TREE     ? "tree"
                                                                                : "auto");
    printf("\nOutput:\n  Directory: %s\n", d->output_dir);
    if (d->checkpoint) printf("  checkpoints: output/checkpoint_NNNNNN.bin\n");
    printf("==============================================\n\n");
}

}  // extern "C"
