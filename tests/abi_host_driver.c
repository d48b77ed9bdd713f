/*
 * Host-only driver of the C ABI's argument paths, built with AddressSanitizer and UBSan
 * (`make -C janus-sim_amd asan`, run by tests/test_sanitize.py): every call below either rejects
 * its arguments before touching a device or fails cleanly when no device is present, and the
 * sanitizers check the host code on those paths (parameter checks, the YAML helper, context
 * teardown after a failed create, snapshot header parsing of bad and truncated files).
 * Prints one line per check and "abi_host_driver: ok" at the end; exits 1 on a wrong return code.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/janus_treepm.h"

static int failures = 0;
#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            fprintf(stderr, "abi_host_driver: FAILED %s (line %d)\n", #cond, __LINE__); \
            ++failures;                                                      \
        }                                                                    \
    } while (0)

static void bad_params(void) {
    jtp_params p;
    jtp_ctx* h = (jtp_ctx*)0x1;
    const int nfields = 14;
    for (int f = 0; f < nfields; ++f) {
        jtp_default_params(&p, 250.0, 64);
        switch (f) {
            case 0: p.n_cells = 48; break;
            case 1: p.box_size = -1.0; break;
            case 2: p.r_cut = 60.0; break;
            case 3: p.theta = NAN; break;
            case 4: p.softening = 0.0; break;
            case 5: p.leaf_size = 17; break;
            case 6: p.group_size = 0; break;
            case 7: p.precision = 7; break;
            case 8: p.split_mode = -1; break;
            case 9: p.gradient = 9; break;
            case 10: p.k_min = -3; break;
            case 11: p.theta_cut = INFINITY; break;
            case 12: p.fft_tile = 13; p.precision = JTP_FP64; break;
            case 13: p.interlace = 5; break;
        }
        h = (jtp_ctx*)0x1;
        CHECK(jtp_create(&p, &h) == JTP_ERR_ARG);
        CHECK(h == NULL);
    }
    CHECK(jtp_create(NULL, &h) == JTP_ERR_ARG);
    CHECK(jtp_create(&p, NULL) == JTP_ERR_ARG);
    printf("bad parameters rejected\n");
}

static void config_fields(void) {
    jtp_config_fields f;
    jtp_params p;
    memset(&f, 0, sizeof(f));
    f.box_size_mpc = 500.0;
    f.n_cells = 128;
    f.n_particles = 2000000;
    f.theta = -1.0;
    CHECK(jtp_params_from_config(&f, &p) == JTP_OK);
    CHECK(p.r_cut == 2.0 * 500.0 / 128 && fabs(p.r_s - p.r_cut / 3.0) < 1e-15 && p.theta == 0.7);
    f.n_particles = 0;
    CHECK(jtp_params_from_config(&f, &p) == JTP_ERR_ARG);
    f.softening_mpc = 0.1;
    f.split_scale_cells = 40.0;  /* r_cut beyond L / 5 */
    CHECK(jtp_params_from_config(&f, &p) == JTP_ERR_ARG);
    f.n_cells = 0;
    CHECK(jtp_params_from_config(&f, &p) == JTP_ERR_ARG);
    CHECK(jtp_params_from_config(NULL, &p) == JTP_ERR_ARG);
    CHECK(jtp_params_from_config(&f, NULL) == JTP_ERR_ARG);
    printf("config fields checked\n");
}

static void null_handles(void) {
    jtp_stats st;
    size_t sz = 0;
    double out3[3];
    double w = 0;
    jtp_diag dg;
    CHECK(jtp_destroy(NULL) == JTP_OK);
    CHECK(strcmp(jtp_last_error(NULL), "null context") == 0);
    CHECK(jtp_version() != NULL && strlen(jtp_version()) > 0);
    CHECK(jtp_set_species_mass(NULL, 1.0, 1.0) == JTP_ERR_ARG);
    CHECK(jtp_compute_forces(NULL, 10, NULL, NULL, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_step(NULL, 10, NULL, NULL, NULL, NULL, 0, 0.1, 0, 0, 1, NULL) == JTP_ERR_ARG);
    CHECK(jtp_get_stats(NULL, &st) == JTP_ERR_ARG);
    CHECK(jtp_stage_size(NULL, JTP_STAGE_KEYS, &sz) == JTP_ERR_ARG);
    CHECK(jtp_stage_copy(NULL, JTP_STAGE_KEYS, out3, sizeof(out3)) == JTP_ERR_ARG);
    CHECK(jtp_pm_solve(NULL, NULL, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_diagnostics(NULL, 0, NULL, NULL, &dg, NULL) == JTP_ERR_ARG);
    CHECK(jtp_potential_energy(NULL, 0, NULL, NULL, &w, NULL) == JTP_ERR_ARG);
    CHECK(jtp_segregation(NULL, 0, NULL, NULL, 8, out3, NULL) == JTP_ERR_ARG);
    CHECK(jtp_pos_to_fixed_f64(NULL, 1, NULL, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_pos_to_fixed_f32(NULL, 1, NULL, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_morton_keys_f32(-1, NULL, 1.0f, 1.0f, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_morton30_keys_f32(5, NULL, 1.0f, 1.0f, NULL, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_morton30_keys_f32((int64_t)1 << 31, NULL, 1.0f, 1.0f, NULL, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_generate_ics(NULL, NULL, NULL, NULL, 0, NULL, NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_ensemble_destroy(NULL) == JTP_OK);
    CHECK(strcmp(jtp_ensemble_last_error(NULL), "null ensemble") == 0);
    CHECK(jtp_ensemble_compute_forces(NULL, NULL, NULL, NULL, NULL, 0, NULL) == JTP_ERR_ARG);
    CHECK(jtp_ensemble_get_stats(NULL, -1, &st) == JTP_ERR_ARG);
    CHECK(jtp_device_count(NULL) == JTP_ERR_ARG);
    CHECK(jtp_device_memory(NULL, NULL) == JTP_ERR_ARG);
    CHECK(jtp_copy_d2d(NULL, NULL, 16) == JTP_ERR_ARG);
    jtp_free(NULL);
    printf("null handles rejected\n");
}

static void ensemble_args(void) {
    jtp_params p[2];
    double mp[2] = {1.0, 1.0}, mm[2] = {19.0, 19.0};
    jtp_ensemble* e = (jtp_ensemble*)0x1;
    jtp_default_params(&p[0], 100.0, 32);
    jtp_default_params(&p[1], 100.0, 48); /* invalid mesh */
    CHECK(jtp_ensemble_create(0, p, mp, mm, 0, &e) == JTP_ERR_ARG);
    CHECK(jtp_ensemble_create(2, p, mp, mm, 99, &e) == JTP_ERR_ARG);
    CHECK(jtp_ensemble_create(2, NULL, mp, mm, 0, &e) == JTP_ERR_ARG);
    e = (jtp_ensemble*)0x1;
    CHECK(jtp_ensemble_create(2, p, mp, mm, 0, &e) != JTP_OK); /* the second box is invalid */
    CHECK(e == NULL);
    printf("ensemble arguments rejected\n");
}

/* A valid create without a device fails in the device part; its partial teardown must be clean. */
static void create_without_device(void) {
    int n = -1;
    jtp_params p;
    jtp_ctx* h = NULL;
    int rc = jtp_device_count(&n);
    if (rc == JTP_OK && n > 0) {
        printf("device present: create-without-device path skipped\n");
        return;
    }
    jtp_default_params(&p, 100.0, 32);
    rc = jtp_create(&p, &h);
    CHECK(rc != JTP_OK && h == NULL);
    CHECK(jtp_alloc(64) == NULL);
    printf("create without a device fails cleanly (rc %d)\n", rc);
}

static void write_file(const char* path, const void* data, size_t n) {
    FILE* f = fopen(path, "wb");
    if (!f) { perror(path); exit(2); }
    fwrite(data, 1, n, f);
    fclose(f);
}

static void snapshot_headers(const char* dir) {
    char path[4096];
    jtp_snap_meta m;
    unsigned char buf[512];
    jtp_snap_meta_default(&m, JTP_SNAP_V3, 250.0);
    CHECK(m.format == JTP_SNAP_V3 && m.box_size == 250.0 && m.velocity_scale == 1.0);
    snprintf(path, sizeof(path), "%s/does_not_exist.bin", dir);
    CHECK(jtp_snapshot_read_meta(path, JTP_SNAP_V3, &m) != JTP_OK);
    CHECK(strlen(jtp_io_last_error()) > 0);
    /* truncated v3 header */
    memset(buf, 0, sizeof(buf));
    memcpy(buf, "JANUSV3\n", 8);
    snprintf(path, sizeof(path), "%s/truncated_v3.bin", dir);
    write_file(path, buf, 100);
    CHECK(jtp_snapshot_read_meta(path, JTP_SNAP_V3, &m) != JTP_OK);
    /* wrong magic, full length */
    memcpy(buf, "NOTJANUS", 8);
    snprintf(path, sizeof(path), "%s/bad_magic_v3.bin", dir);
    write_file(path, buf, sizeof(buf));
    CHECK(jtp_snapshot_read_meta(path, JTP_SNAP_V3, &m) != JTP_OK);
    /* production header shorter than 16 bytes */
    snprintf(path, sizeof(path), "%s/short_production.bin", dir);
    write_file(path, buf, 7);
    CHECK(jtp_snapshot_read_meta(path, JTP_SNAP_PRODUCTION, &m) != JTP_OK);
    CHECK(jtp_snapshot_read_meta(path, 2, &m) != JTP_OK);
    CHECK(jtp_snapshot_read_meta(NULL, JTP_SNAP_V3, &m) != JTP_OK);
    CHECK(jtp_snapshot_write(NULL, NULL, NULL, NULL, 0, NULL, NULL) != JTP_OK);
    printf("snapshot headers checked\n");
}

static void ic_defaults(void) {
    jtp_ic_params ic;
    jtp_ic_default_params(&ic);
    CHECK(ic.box_size == 50.0 && ic.n_grid == 128 && ic.seed == 42 && ic.mu == 19.0);
    jtp_ic_default_params(NULL);
    jtp_default_params(NULL, 1.0, 16);
    jtp_snap_meta_default(NULL, JTP_SNAP_V3, 1.0);
    printf("defaults checked\n");
}

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : ".";
    bad_params();
    config_fields();
    null_handles();
    ensemble_args();
    create_without_device();
    snapshot_headers(dir);
    ic_defaults();
    if (failures) {
        fprintf(stderr, "abi_host_driver: %d check(s) failed\n", failures);
        return 1;
    }
    printf("abi_host_driver: ok\n");
    return 0;
}
