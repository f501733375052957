// Parameter derivation: the reference's set_cpu_param / set_gpu_param /
// check_dt (single_gpu_sph_src/lib/init.cu:153-251, lib/check.cu:3-18),
// with the 2-D normalisation, explicit sound speed, CFL time step and EOS
// exponent added as case parameters.
#include <cmath>
#include <cstring>

#include "zsph.h"
#include "host_internal.hpp"

extern "C" void zsph_json_defaults(zsph_json *j) {
    std::memset(j, 0, sizeof(*j));
    j->g = -9.81f;
    j->rho0 = 1000.0f;
    j->dx = 0.01f;
    j->h_factor = 1.3f;
    j->r_factor = 2.0f;
    j->eta_factor = 0.1f;
    j->cs_factor = 10.0f;
    j->delta = 0.1f;
    j->alpha = 0.02f;
    j->grid_layer_factor = 1.0f;
    j->grid_size_factor = 1.0f;
    j->dt = 0.0f;
    j->start_step = 0;
    j->current_step = 0;
    j->end_step = 0;
    j->output_step = 100;
    j->thread_num = 1;
    j->gpu_id = 0;
    std::strcpy(j->input_path, ".");
    std::strcpy(j->output_path, ".");
    j->dim = 3;
    j->gamma = 1.0f;
    j->c0 = 0.0f;
    j->cfl = 0.0f;
    j->diffusion = ZSPH_DIFF_ANTUONO;
    j->output_fields = ZSPH_F_POS | ZSPH_F_VEL | ZSPH_F_RHOP | ZSPH_F_TYPE;
    j->binary = 1;
}

// lib/init.cu:153-219.  All arithmetic is single precision in the same
// order as the reference so the grid (and therefore every cell hash)
// matches it for the same case file.
extern "C" int zsph_set_cpu_param(zsph_cpu_param *p, const zsph_json *j) {
    if (!p || !j) return ZSPH_E_ARG;
    std::memset(p, 0, sizeof(*p));
    if (j->dim != 2 && j->dim != 3) return ZSPH_E_ARG;
    if (!(j->dx > 0.0f) || !(j->h_factor > 0.0f) || !(j->grid_size_factor >= 1.0f) ||
        !(j->rho0 > 0.0f) || j->ptc_num < 0)
        return ZSPH_E_ARG;
    p->dim = j->dim;
    p->g = j->g;
    const float dx = j->dx;
    p->m = (j->dim == 3) ? j->rho0 * dx * dx * dx : j->rho0 * dx * dx;
    p->rho0 = j->rho0;
    p->dx = dx;
    p->h = j->h_factor * dx;
    p->r = j->r_factor * j->h_factor * dx;
    p->eta = j->eta_factor * j->h_factor * dx;
    if (j->c0 > 0.0f)
        p->cs = j->c0;
    else
        p->cs = j->cs_factor * std::sqrt(std::fabs(j->g) * (j->zmax - j->zmin));
    p->delta = j->delta;
    p->alpha = j->alpha;
    // Wendland C2 normalisation: 21/(16 pi h^3) in 3-D, 7/(4 pi h^2) in 2-D.
    const double hd = (double)p->h;
    p->adh = (j->dim == 3) ? (float)(21.0 / (16.0 * M_PI * hd * hd * hd))
                           : (float)(7.0 / (4.0 * M_PI * hd * hd));

    p->xmin = j->xmin; p->xmax = j->xmax;
    p->ymin = j->ymin; p->ymax = j->ymax;
    p->zmin = j->zmin; p->zmax = j->zmax;

    p->grid_size = j->r_factor * j->h_factor * dx / j->grid_size_factor;
    p->grid_factor = (int)j->grid_size_factor;
    // The cell walk covers grid_factor cells each side: it must reach 2h.
    if (!(p->grid_size * (float)p->grid_factor >= 2.0f * p->h)) return ZSPH_E_ARG;
    const float pad = p->grid_size * j->grid_layer_factor;
    p->grid_xmin = j->xmin - pad; p->grid_xmax = j->xmax + pad;
    p->grid_zmin = j->zmin - pad; p->grid_zmax = j->zmax + pad;
    p->grid_xdim = (int)((p->grid_xmax - p->grid_xmin) / p->grid_size);
    p->grid_zdim = (int)((p->grid_zmax - p->grid_zmin) / p->grid_size);
    if (j->dim == 3) {
        p->grid_ymin = j->ymin - pad; p->grid_ymax = j->ymax + pad;
        p->grid_ydim = (int)((p->grid_ymax - p->grid_ymin) / p->grid_size);
    } else {
        // 2-D (x-z plane): a single cell row centred on y = 0.
        p->grid_ymin = j->ymin - 0.5f * p->grid_size;
        p->grid_ymax = p->grid_ymin + p->grid_size;
        p->grid_ydim = 1;
    }
    if (p->grid_xdim < 1 || p->grid_ydim < 1 || p->grid_zdim < 1) return ZSPH_E_ARG;
    const long long G = (long long)p->grid_xdim * p->grid_ydim * p->grid_zdim;
    if (G > (1ll << 30)) return ZSPH_E_ARG;
    p->grid_num = (int)G;
    p->grid_hash_min = 0;
    p->grid_hash_max = p->grid_num - 1;

    p->ptc_num = j->ptc_num;
    p->water_ptc_num = j->water_ptc_num;
    p->air_ptc_num = j->air_ptc_num;
    p->wall_ptc_num = j->wall_ptc_num;
    p->rigid_ptc_num = j->rigid_ptc_num;
    p->thread_num = j->thread_num;
    p->gpu_id = j->gpu_id;
    std::memcpy(p->input_path, j->input_path, sizeof(p->input_path));
    std::memcpy(p->output_path, j->output_path, sizeof(p->output_path));
    std::memcpy(p->git_hash, j->git_hash, sizeof(p->git_hash));

    p->dt = (j->cfl > 0.0f) ? j->cfl * p->h / p->cs : j->dt;
    p->start_step = j->start_step;
    p->current_step = j->current_step;
    p->end_step = j->end_step;
    p->output_step = j->output_step;
    p->gamma = (j->gamma > 0.0f) ? j->gamma : 1.0f;
    p->diffusion = j->diffusion;
    p->output_fields = j->output_fields;
    p->binary = j->binary;
    if (p->diffusion != ZSPH_DIFF_ANTUONO && p->diffusion != ZSPH_DIFF_MC) return ZSPH_E_ARG;
    return zsph_check_dt(p);
}

// lib/check.cu:3-18: dt <= 0.25 h / sqrt(|g| (zmax - zmin)).
int zsph_check_dt(const zsph_cpu_param *p) {
    if (!(p->dt > 0.0f)) return ZSPH_E_DT;
    const float H = p->zmax - p->zmin;
    const float gH = std::fabs(p->g) * H;
    if (gH > 0.0f) {
        const float dt_max = 0.25f * p->h / std::sqrt(gH);
        if (p->dt > dt_max) return ZSPH_E_DT;
    }
    return ZSPH_OK;
}

// lib/init.cu:221-251.
extern "C" void zsph_set_gpu_param(zsph_params *g, const zsph_cpu_param *c) {
    std::memset(g, 0, sizeof(*g));
    g->half_dt = c->dt / 2.0f;
    g->h = c->h;
    g->g = c->g;
    g->m = c->m;
    g->rho0 = c->rho0;
    g->adh = c->adh;
    g->h2 = c->h * c->h;
    g->eta2 = c->eta * c->eta;
    g->cs2 = c->cs * c->cs;
    g->alpha = c->alpha;
    g->delta_h_cs = c->delta * c->h * c->cs;
    g->h_cs_rho_m = c->h * c->cs * c->rho0 * c->m;
    g->grid_factor = c->grid_factor;
    g->grid_size = c->grid_size;
    g->grid_xmin = c->grid_xmin;
    g->grid_ymin = c->grid_ymin;
    g->grid_zmin = c->grid_zmin;
    g->grid_xdim = c->grid_xdim;
    g->grid_ydim = c->grid_ydim;
    g->grid_zdim = c->grid_zdim;
    g->grid_hash_min = c->grid_hash_min;
    g->grid_hash_max = c->grid_hash_max;
    g->ptc_num = c->ptc_num;
    g->dim = c->dim;
    g->gamma = c->gamma;
    g->diffusion = c->diffusion;
    g->rc2 = 4.0f * g->h2;
}

extern "C" size_t zsph_sizeof(int which) {
    switch (which) {
    case 0: return sizeof(zsph_json);
    case 1: return sizeof(zsph_cpu_param);
    case 2: return sizeof(zsph_params);
    case 3: return sizeof(zsph_host_state);
    case 4: return sizeof(zsph_dambreak);
    case 5: return sizeof(zsph_rigid);
    case 6: return sizeof(zsph_diag);
    case 7: return sizeof(zsph_tuning);
    case 8: return sizeof(zsph_piston);
    case 9: return sizeof(zsph_wavetank);
    case 10: return sizeof(zsph_closure);
    case 11: return sizeof(zsph_checkpoint_info);
    default: return 0;
    }
}
