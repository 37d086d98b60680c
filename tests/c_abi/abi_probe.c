/* A plain C caller of libzmpm.so, compiled by gcc against include/zmpm.h only.
 *
 * It prints the size and field offsets of every struct the header declares, as
 * "<struct> <field> <offset>" lines ("<struct> sizeof <bytes>" for the size), so a
 * test can compare them with the Python binding's ctypes layouts, and then makes
 * a few calls through the ABI: context creation (ZMPM_ERR_NODEVICE without a
 * GPU), the last-error string, and the host-side view check. */
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "zmpm.h"

#define SZ(S) printf(#S " sizeof %zu\n", sizeof(S))
#define OFF(S, f) printf(#S " " #f " %zu\n", offsetof(S, f))

static void layouts(void) {
    SZ(zmpm_material);
    OFF(zmpm_material, model); OFF(zmpm_material, E); OFF(zmpm_material, nu); OFF(zmpm_material, rho);
    OFF(zmpm_material, vol); OFF(zmpm_material, bulk); OFF(zmpm_material, scene); OFF(zmpm_material, plasticity);
    OFF(zmpm_material, friction_angle); OFF(zmpm_material, yield_stress); OFF(zmpm_material, flip_beta);
    OFF(zmpm_material, nacc_beta); OFF(zmpm_material, nacc_xi); OFF(zmpm_material, aniso);
    OFF(zmpm_material, aniso_strength);

    SZ(zmpm_config);
    OFF(zmpm_config, dx); OFF(zmpm_config, n_cells); OFF(zmpm_config, capacity); OFF(zmpm_config, device);

    SZ(zmpm_scene);
    OFF(zmpm_scene, dx); OFF(zmpm_scene, n_cells); OFF(zmpm_scene, dt); OFF(zmpm_scene, gravity);
    OFF(zmpm_scene, boundary); OFF(zmpm_scene, pad);

    SZ(zmpm_step_params);
    OFF(zmpm_step_params, dt); OFF(zmpm_step_params, gravity); OFF(zmpm_step_params, boundary);
    OFF(zmpm_step_params, pad);

    SZ(zmpm_collider);
    OFF(zmpm_collider, type); OFF(zmpm_collider, shape); OFF(zmpm_collider, scene); OFF(zmpm_collider, dims);
    OFF(zmpm_collider, shape_params); OFF(zmpm_collider, sdf); OFF(zmpm_collider, origin);
    OFF(zmpm_collider, spacing); OFF(zmpm_collider, translation); OFF(zmpm_collider, translation_rate);
    OFF(zmpm_collider, scale); OFF(zmpm_collider, scale_rate); OFF(zmpm_collider, ypr_degrees);
    OFF(zmpm_collider, angular_velocity); OFF(zmpm_collider, leaf_origins); OFF(zmpm_collider, nleaves);
    OFF(zmpm_collider, reserved);

    SZ(zmpm_particles);
    OFF(zmpm_particles, n); OFF(zmpm_particles, layout); OFF(zmpm_particles, tile_channels);
    OFF(zmpm_particles, x); OFF(zmpm_particles, x_chn); OFF(zmpm_particles, v); OFF(zmpm_particles, v_chn);
    OFF(zmpm_particles, C); OFF(zmpm_particles, C_chn); OFF(zmpm_particles, F); OFF(zmpm_particles, F_chn);
    OFF(zmpm_particles, J); OFF(zmpm_particles, J_chn); OFF(zmpm_particles, material);
    OFF(zmpm_particles, beta); OFF(zmpm_particles, beta_chn); OFF(zmpm_particles, logJp);
    OFF(zmpm_particles, logJp_chn); OFF(zmpm_particles, mu); OFF(zmpm_particles, mu_chn);
    OFF(zmpm_particles, lam); OFF(zmpm_particles, lam_chn); OFF(zmpm_particles, a); OFF(zmpm_particles, a_chn);
    OFF(zmpm_particles, m); OFF(zmpm_particles, m_chn); OFF(zmpm_particles, vol); OFF(zmpm_particles, vol_chn);
    OFF(zmpm_particles, default_material);

    SZ(zmpm_boundary_particles);
    OFF(zmpm_boundary_particles, n); OFF(zmpm_boundary_particles, x); OFF(zmpm_boundary_particles, v);
    OFF(zmpm_boundary_particles, m); OFF(zmpm_boundary_particles, nrm); OFF(zmpm_boundary_particles, scene);

    SZ(zmpm_adjoint);
    OFF(zmpm_adjoint, n); OFF(zmpm_adjoint, x); OFF(zmpm_adjoint, v); OFF(zmpm_adjoint, C);
    OFF(zmpm_adjoint, F); OFF(zmpm_adjoint, mu); OFF(zmpm_adjoint, lam);
}

int main(void) {
    layouts();
    /* a context: ZMPM_ERR_NODEVICE where no gfx950 device is visible, ZMPM_OK on one */
    zmpm_ctx *ctx = NULL;
    zmpm_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.dx = 1.0f / 64.0f;
    cfg.n_cells = 64;
    cfg.capacity = 1000;
    cfg.device = 0;
    int32_t rc = zmpm_create(&ctx, &cfg);
    printf("call create %d\n", (int)rc);
    if (rc == ZMPM_OK) {
        printf("call destroy %d\n", (int)zmpm_destroy(ctx));
    } else {
        printf("error %s\n", zmpm_last_error());
    }
    /* argument errors: no device is touched */
    rc = zmpm_create(NULL, &cfg);
    printf("call create_null %d\n", (int)rc);
    float dummy[64];
    zmpm_particles view;
    memset(&view, 0, sizeof(view));
    view.n = 10;
    view.layout = ZMPM_LAYOUT_AOSOA32;
    view.tile_channels = 20;
    view.x = dummy;
    view.x_chn = 1;
    view.J = dummy;
    view.J_chn = 19;
    printf("call check_ok %d\n", (int)zmpm_check_particles(&view));
    view.F = dummy;
    view.F_chn = 19; /* 9 channels from 19 in a 20-channel tile */
    rc = zmpm_check_particles(&view);
    printf("call check_bad %d\n", (int)rc);
    printf("error %s\n", zmpm_last_error());
    return 0;
}
